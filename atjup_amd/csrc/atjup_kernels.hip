// HIP kernels of the ATJUP time step for gfx950 (MI355X), FP64 drop-in and FP32 form.
//
// Layout: every field is one contiguous array [i][j][k] with the k row padded to `pitch`
// elements (128-byte aligned rows), k (phi) fastest, so a 64-wide wavefront reads one row
// segment per load. The stencil kernels walk a 1-D grid whose block index is remapped so that
// each of the 8 XCDs owns a contiguous slab of i-planes: the vertical (i +- 1) and meridional
// (j +- 1) neighbours a block reads were brought in by blocks of the same XCD and hit its L2.
//
// All FP64 kernels keep the reference's expression trees (the file is built with
// -ffp-contract=off), so that a cell update reproduces the CPU oracle bit for bit whenever its
// inputs agree.
#include "atjup_device.h"

#include <algorithm>
#include <cmath>

namespace atjup {

// ---------------------------------------------------------------------------------------------
// computeBuoyancyRefLevel (RungeKutta_Jup_Turb.cpp:47-69): per level i the sin(theta)-weighted
// mean of g p_stat / (r_mix R_mix t t_ref) over fluid cells with t > 0. Two deterministic passes:
// NCH chunks of (j, k) per level reduce to partial sums in a fixed order, then one thread per
// level adds its partials in chunk order.
// ---------------------------------------------------------------------------------------------
constexpr int BRL_NCH = 32;
constexpr int BRL_BLOCK = 256;

template<typename T>
__global__ __launch_bounds__(BRL_BLOCK) void k_brl_partial(const T* __restrict__ t, const T* __restrict__ ps,
        const uint8_t* __restrict__ flags, const double* __restrict__ sin_the, acc_t<T>* __restrict__ part,
        int jm, int km, int pitch, double g, double rR, double t_ref){
    using A = acc_t<T>;
    const int i = blockIdx.y, ch = blockIdx.x;
    const int ncell = jm * km;
    const int per = (ncell + BRL_NCH - 1) / BRL_NCH;
    const int c0 = ch * per, c1 = min(ncell, c0 + per);
    A s = 0.0, w = 0.0;
    // every load unconditional and the skipped cells' terms left out by selection, so that a
    // thread's loads are in flight together; the sums run in the same order as a skipping loop
    for(int c = c0 + threadIdx.x; c < c1; c += BRL_BLOCK){
        const int j = c / km, k = c - j * km;
        const size_t q = ((size_t)i * jm + j) * pitch + k;
        const bool solid = flags[q] & FL_SELF;
        const A tq = (A)t[q];
        const A pq = (A)ps[q];
        const double wgt = sin_the[j];
        const bool warm = tq > 0.0;
        const A b = g * pq / (rR * (warm ? tq : A(1.0)) * t_ref);
        if(!solid && warm && isfinite(b)){
            s += wgt * b;
            w += wgt;
        }
    }
    __shared__ A sh[2][BRL_BLOCK];
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = w;
    __syncthreads();
    for(int o = BRL_BLOCK / 2; o > 0; o >>= 1){
        if((int)threadIdx.x < o){
            sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if(threadIdx.x == 0){
        part[((size_t)i * BRL_NCH + ch) * 2 + 0] = sh[0][0];
        part[((size_t)i * BRL_NCH + ch) * 2 + 1] = sh[1][0];
    }
}

template<typename A>
__global__ void k_brl_final(const A* __restrict__ part, A* __restrict__ brl, int im){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= im) return;
    A s = 0.0, w = 0.0;
    for(int ch = 0; ch < BRL_NCH; ch++){
        s += part[((size_t)i * BRL_NCH + ch) * 2 + 0];
        w += part[((size_t)i * BRL_NCH + ch) * 2 + 1];
    }
    brl[i] = (w > 0.0) ? s / w : A(0.0);
}

// ---------------------------------------------------------------------------------------------
// computeHydrostaticPressure, METRIC_SPLIT + CORIOLIS_SPLIT without HYDRO_SPLIT and p_baro
// (RungeKutta_Jup_Turb.cpp:119-218): one thread per (j, k) column, trapezoid scan over the
// column's fluid run from its foot, zero below and above it.
// ---------------------------------------------------------------------------------------------
template<typename T>
__global__ __launch_bounds__(256) void k_hydrostatic(const T* __restrict__ v, const T* __restrict__ w,
        const int2* __restrict__ colr, const double* __restrict__ rad, const double* __restrict__ cor_j,
        T* __restrict__ ph, int im, int jm, int km, int pitch, size_t plane, double dr){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    // the column's fluid run from the bottom, fixed by the mask (k_col_range)
    const int2 run = colr[(size_t)j * km + k];
    const int i0 = run.x, i1 = run.y;
    for(int i = 0; i < i0 && i < im; i++) ph[q0 + i * plane] = T(0);
    if(i0 >= im) return;
    for(int i = i1 + 1; i < im; i++) ph[q0 + i * plane] = T(0);
    // the column integral is formed in FP64 in both forms
    using A = acc_t<T>;
    const double cor_ = cor_j[j];
    auto f_ = [&](int i) -> A {
        const A vv = (A)v[q0 + i * plane], ww = (A)w[q0 + i * plane];
        return 0.0 + ((vv * vv + ww * ww) / rad[i] + cor_ * ww) / 1.0;
    };
    A prev = 0.0;
    ph[q0 + i0 * plane] = (T)prev;
    A fprev = f_(i0);
    // the integrands of eight levels formed ahead of the in-order sum (their loads in flight
    // together; a level past i1 reads i1 again and is not used)
    for(int base = i0 + 1; base <= i1; base += 8){
        A fv[8];
        #pragma unroll
        for(int u = 0; u < 8; u++) fv[u] = f_(min(base + u, i1));
        #pragma unroll
        for(int u = 0; u < 8; u++){
            const int i = base + u;
            if(i > i1) break;
            prev = prev + 0.5 * (fprev + fv[u]) * dr;
            ph[q0 + i * plane] = (T)prev;
            fprev = fv[u];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// RK4 stage with the RHS of the four dynamical fields t, u, v, w (RHSJup, RHS_Jup_Turb.cpp:31-1457,
// under the dry dynamics profile: no species, closure, radiation or precipitation; nue = 0).
// One thread per cell; the stage accumulation and the next stage input (pass B,
// RungeKutta_Jup_Turb.cpp:420-490) are fused into the same kernel, double-buffered so that no
// cell is read after it is written. Stage 3 writes the final assembly (:494-568) in place into
// the time-level-n buffer and the aux_* fields the next pressure solve reads.
// ---------------------------------------------------------------------------------------------
template<typename T>
struct StageArgs {
    const T* in[4];
    T* out[4];
    const T* nl[4];
    T* acc[4];
    const T *pd, *ph, *ps, *rho, *wn;
    T *aux_u, *aux_v, *aux_w;
    const uint8_t* flags;
    const GeoIJ<geo_t<T>>* geo;
    const acc_t<T>* brl;
    int im, jm, km, pitch, nkb, njb, nb;
    size_t plane;
    int stage, copy_boundary;
    T c_in, wgt, dt, one_sixth;
    T inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
    double cbuoy, g, rR, t_ref_d;          // the buoyancy anomaly is formed in FP64 in both forms
    T cep, to_nd, r_mix, c_m2o, c_p2o, Coriolis, centrifugal, t_ref, kt, visc0;
    float* lo[4];                          // FP32-refined form: the low words of t, u, v, w (else null)
    double dt_d;
    const double* chi;                     // the smooth drag region's indicator (else null)
    double inv_tau;                        // its inverse drag time
};

// The FP32-refined form carries t, u, v, w as an FP32 pair hi + lo. The last RK stage adds the
// step's increment to hi + lo in FP64 and splits the sum again, so that an increment of a few
// ulps of the field is not rounded the same way step after step (the stencils read hi only).
__device__ __forceinline__ void split_store(float* hi, float* lo, size_t q, double v){
    const float h = (float)v;
    hi[q] = h;
    lo[q] = (float)(v - (double)h);
}

constexpr int ST_BK = 64, ST_BJ = 4;

template<typename T>
__global__ __launch_bounds__(ST_BK * ST_BJ) void k_stage_dry(const StageArgs<T> a,
        const GeoIJ<geo_t<T>>* __restrict__ geo){
    const int b = xcd_remap(blockIdx.x, a.nb);
    const int kb = b % a.nkb;
    const int rest = b / a.nkb;
    const int jb = rest % a.njb;
    const int i = rest / a.njb;
    const int k = kb * ST_BK + threadIdx.x;
    const int j = jb * ST_BJ + threadIdx.y;
    if(k >= a.km || j >= a.jm) return;
    const size_t q = ((size_t)i * a.jm + j) * a.pitch + k;
    const uint8_t fl = a.flags[q];
    const bool interior = i >= 1 && i <= a.im - 2 && j >= 3 && j <= a.jm - 4 && k <= a.km - 2 && !(fl & FL_SELF);
    if(!interior){
        if(a.copy_boundary){
            #pragma unroll
            for(int f = 0; f < 4; f++) a.out[f][q] = a.in[f][q];
        }
        return;
    }
    const size_t P = a.plane, R = (size_t)a.pitch;
    const int kw = (k == 0) ? a.km - 2 : k - 1;
    const int ke = (k == a.km - 2) ? 0 : k + 1;
    const size_t qkw = q - k + kw, qke = q - k + ke;
    const GeoIJ<geo_t<T>> G = geo[wave_uniform(i * a.jm + j)];     // a wavefront is one (i, j) row

    T c_[4], d1r[4], d2r[4], d1t[4], d2t[4], d1p[4], d2p[4];
    #pragma unroll
    for(int f = 0; f < 4; f++){
        const T* X = a.in[f];
        const T xc = X[q], xip = X[q + P], xim = X[q - P], xjp = X[q + R], xjm = X[q - R];
        const T xkp = X[qke], xkm = X[qkw];
        c_[f] = xc;
        d1r[f] = (xip - xim) * a.inv_2dr;
        d2r[f] = (xip - T(2) * xc + xim) * a.inv_dr2;
        d1t[f] = (xjp - xjm) * a.inv_2dthe;
        d2t[f] = (xjp - T(2) * xc + xjm) * a.inv_dthe2;
        d1p[f] = (xkp - xkm) * a.inv_2dphi;
        d2p[f] = (xkp - T(2) * xc + xkm) * a.inv_dphi2;
    }
    // wall-mirrored pressure gradients (RHS_Jup_Turb.cpp:119-130)
    const T pdc = a.pd[q], phc = a.ph[q];
    const T dpdr   = (((fl & FL_IP) ? pdc : a.pd[q + P]) - ((fl & FL_IM) ? pdc : a.pd[q - P])) * a.inv_2dr;
    const T dpdthe = (((fl & FL_JP) ? pdc : a.pd[q + R]) - ((fl & FL_JM) ? pdc : a.pd[q - R])) * a.inv_2dthe;
    const T dpdphi = (((fl & FL_KP) ? pdc : a.pd[qke]) - ((fl & FL_KM) ? pdc : a.pd[qkw])) * a.inv_2dphi;
    const T dphdthe = (((fl & FL_JP) ? phc : a.ph[q + R]) - ((fl & FL_JM) ? phc : a.ph[q - R])) * a.inv_2dthe;
    const T dphdphi = (((fl & FL_KP) ? phc : a.ph[qke]) - ((fl & FL_KM) ? phc : a.ph[qkw])) * a.inv_2dphi;

    const T tq = c_[0], u = c_[1], v = c_[2], w = c_[3];
    const T costhe = G.costhe, st = G.sinthe_true;
    const T Coriolis_the = a.c_m2o * costhe * w;
    const T Coriolis_phi = a.c_p2o * (+costhe * v + st * u);
    const T v_invrm = v * G.inv_rm;
    const T w_invrs = w * G.inv_rmsinthe;

    T dpstatdr = (a.ps[q + P] - a.ps[q - P]) * a.inv_2dr * a.to_nd;
    {
        const T rho_l = a.rho[q];
        if(rho_l > T(0) && isfinite(rho_l)) dpstatdr *= a.r_mix / rho_l;
    }
    const T pressure_t = a.cep * (u * (dpdr + dpstatdr) + v_invrm * dpdthe + w_invrs * dpdphi);
    const T transport_t = u * d1r[0] + v_invrm * d1t[0] + w_invrs * d1p[0];
    const T transport_u = u * d1r[1] + v_invrm * d1t[1] + w_invrs * d1p[1]
        - T(0) * (v * v + w * w) * G.inv_rm;
    const T transport_v = u * d1r[2] + v_invrm * d1t[2] + w_invrs * d1p[2]
        + (u * v - w * w * G.cotanthe) * G.inv_rm;
    const T transport_w = u * d1r[3] + v_invrm * d1t[3] + w_invrs * d1p[3]
        + (w * u + v * w * G.cotanthe) * G.inv_rm;
    const T two_inv_rm = T(2) * G.inv_rm;
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

    const T rhs_t = + pressure_t - transport_t + diffusion_t * a.kt;
    const T buoyancy_u = (tq > T(0))
        ? (T)(a.cbuoy * (a.g * (acc_t<T>)a.ps[q] / (a.rR * (acc_t<T>)tq * a.t_ref_d) - a.brl[i])) : T(0);
    const T visc = a.visc0 + a.wn[q];
    T rhs_u = - dpdr + buoyancy_u - transport_u + diffusion_u * visc
        + a.centrifugal * G.cent_rad;
    T rhs_v = - dpdthe_term - dphdthe_term - transport_v + diffusion_v * visc
        - a.Coriolis * Coriolis_the + a.centrifugal * G.cent_the;
    T rhs_w = - dpdphi_term - dphdphi_term - transport_w + diffusion_w * visc
        - a.Coriolis * Coriolis_phi;
    if(a.chi){
        // the smooth drag region, -chi / tau on each velocity component
        const T pen = (T)(a.chi[q] * a.inv_tau);
        rhs_u = rhs_u - pen * u;
        rhs_v = rhs_v - pen * v;
        rhs_w = rhs_w - pen * w;
    }
    const T rhs[4] = { rhs_t, rhs_u, rhs_v, rhs_w };

    if(a.stage == 0){
        #pragma unroll
        for(int f = 0; f < 4; f++){
            a.acc[f][q] = a.wgt * rhs[f];
            a.out[f][q] = c_[f] + a.c_in * rhs[f];       // stage 0 reads the n level itself
        }
    } else if(a.stage < 3){
        #pragma unroll
        for(int f = 0; f < 4; f++){
            a.acc[f][q] = a.acc[f][q] + a.wgt * rhs[f];
            a.out[f][q] = a.nl[f][q] + a.c_in * rhs[f];
        }
    } else {
        #pragma unroll
        for(int f = 0; f < 4; f++){
            const T accf = a.acc[f][q] + a.wgt * rhs[f];
            if constexpr(std::is_same<T, float>::value){
                // the low word loaded without a branch (from the n level where there is none), so
                // that the wavefront does not wait at the branch for its loads in flight
                const T nlv = a.nl[f][q];
                const float lov = (a.lo[f] ? a.lo[f] : (const float*)a.nl[f])[q];
                if(a.lo[f]){
                    split_store(a.out[f], a.lo[f], q, ((double)nlv + (double)lov) + a.dt_d * (double)accf * (1.0 / 6.0));
                    continue;
                }
                a.out[f][q] = nlv + a.dt * accf * a.one_sixth;
                continue;
            }
            a.out[f][q] = a.nl[f][q] + a.dt * accf * a.one_sixth;
        }
        a.aux_u[q] = rhs_u + dpdr;
        a.aux_v[q] = rhs_v + dpdthe_term;
        a.aux_w[q] = rhs_w + dpdphi_term;
    }
}

// ---------------------------------------------------------------------------------------------
// Red-black Poisson solve (PressureSolver.h:145-766).
// ---------------------------------------------------------------------------------------------
// prepareProjectionBoundaries, rigid lid (cJupiterModel.cpp:1842-1864)
template<typename T>
__global__ void k_prep_projection(T* __restrict__ au, T* __restrict__ av, T* __restrict__ aw,
        int im, int jm, int km, int pitch, size_t plane, double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    if(k > km - 2) return;
    const size_t q = (size_t)j * pitch + k;
    const size_t t0 = q, t1 = q + plane, t2 = q + 2 * plane;
    const size_t b0 = q + (size_t)(im - 1) * plane, b1 = q + (size_t)(im - 2) * plane, b2 = q + (size_t)(im - 3) * plane;
    au[t0] = T(0);
    au[b0] = T(0);
    av[t0] = (T)c43 * av[t1] - (T)c13 * av[t2];
    av[b0] = (T)c43 * av[b1] - (T)c13 * av[b2];
    aw[t0] = (T)c43 * aw[t1] - (T)c13 * aw[t2];
    aw[b0] = (T)c43 * aw[b1] - (T)c13 * aw[b2];
}

// The divergence source of every cell of the sweep domain, computed once per call: it depends
// only on aux_* and the mask, which the sweeps do not change. Clamped to +-den * PDYN_CAP with the
// cell's own denominator (PressureSolver.h:370-487).
template<typename T, typename I>
__global__ __launch_bounds__(256) void k_poisson_source(const T* __restrict__ au, const T* __restrict__ av,
        const T* __restrict__ aw, const uint8_t* __restrict__ flags, const uint8_t* __restrict__ msk,
        const PoisIJ<double>* __restrict__ pw, I* __restrict__ src0, I* __restrict__ src1,
        int im, int jm, int km, int pitch, size_t plane, int nb, int hp,
        double inv_2dr, double inv_2dthe, double inv_2dphi, double inv_dthe, double inv_dphi, double cap){
    const int b = xcd_remap(blockIdx.x, nb);
    const int nkb = (km - 1 + 63) / 64;
    const int kb = b % nkb, rest = b / nkb;
    const int j = rest % (jm - 2) + 1, i = rest / (jm - 2) + 1;
    const int k = kb * 64 + threadIdx.x;
    if(k > km - 2 || i > im - 2) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const size_t P = plane, R = (size_t)pitch;
    const uint8_t fl = flags[q];
    const bool sld = fl & FL_SELF;
    const int kw = (k == 0) ? km - 2 : k - 1, ke = (k == km - 2) ? 0 : k + 1;
    const bool skp = fl & FL_KP, skm = fl & FL_KM;
    const bool skp2 = (k < km - 2) ? (msk[q + 2] != 0) : false;
    const bool skm2 = (k > 2) ? (msk[q - 2] != 0) : false;
    I du_dr = 0.0, dv_dthe = 0.0, dw_dphi = 0.0;
    bool r_flag = false, the_flag = false, phi_flag = false;
    if(i < im - 2 && sld && !(fl & FL_IP)){
        du_dr = (-3.0 * (I)au[q] + 4.0 * (I)au[q + P] - (I)au[q + 2 * P]) * inv_2dr;
        r_flag = true;
    }
    if(j > 2 && j < jm - 2){
        const bool air_jp1 = !(fl & FL_JP), air_jm1 = !(fl & FL_JM);
        const bool sjp2 = msk[q + 2 * R] != 0, sjm2 = msk[q - 2 * R] != 0;
        if(sld && air_jp1 && !sjp2){
            dv_dthe = (-3.0 * (I)av[q] + 4.0 * (I)av[q + R] - (I)av[q + 2 * R]) * inv_2dthe;
            the_flag = true;
        }
        if(sld && air_jm1 && !sjm2){
            dv_dthe = -(-3.0 * (I)av[q] + 4.0 * (I)av[q - R] - (I)av[q - 2 * R]) * inv_2dthe;
            the_flag = true;
        }
        if((sld && air_jp1 && sjp2) || (j == jm - 2 && !sld && (fl & FL_JP))){
            dv_dthe = ((I)av[q + R] - (I)av[q]) * inv_dthe;
            the_flag = true;
        }
        if((sld && air_jm1 && sjm2) || (j == 1 && sld && air_jm1)){
            dv_dthe = ((I)av[q - R] - (I)av[q]) * inv_dthe;
            the_flag = true;
        }
    }
    if(k > 2 && k < km - 2){
        const bool air_kp1 = !skp, air_km1 = !skm;
        if(sld && air_kp1 && !skp2){
            dw_dphi = (-3.0 * (I)aw[q] + 4.0 * (I)aw[q + 1] - (I)aw[q + 2]) * inv_2dphi;
            phi_flag = true;
        }
        if(sld && air_km1 && !skm2){
            dw_dphi = -(-3.0 * (I)aw[q] + 4.0 * (I)aw[q - 1] - (I)aw[q - 2]) * inv_2dphi;
            phi_flag = true;
        }
        if((sld && air_kp1 && skp2) || (k == km - 2 && !sld && skp)){
            dw_dphi = ((I)aw[q + 1] - (I)aw[q]) * inv_dphi;
            phi_flag = true;
        }
        if((sld && air_km1 && skm2) || (k == 1 && sld && air_km1)){
            dw_dphi = ((I)aw[q - 1] - (I)aw[q]) * inv_dphi;
            phi_flag = true;
        }
    }
    if(!r_flag) du_dr = ((I)au[q + P] - (I)au[q - P]) * inv_2dr;
    if(!the_flag) dv_dthe = ((I)av[q + R] - (I)av[q - R]) * inv_2dthe;
    if(!phi_flag) dw_dphi = ((I)aw[q - k + ke] - (I)aw[q - k + kw]) * inv_2dphi;
    const PoisIJ<double> W = pw[(size_t)i * jm + j];
    I div_src = du_dr * 1.0 + dv_dthe * W.inv_rm + dw_dphi * W.inv_rmsinthe;
    double den = W.denom;
    if(!sld){
        den = 0.0;
        if(!(fl & FL_IP)) den += W.num1;
        if(!(fl & FL_IM)) den += W.num1;
        if(!(fl & FL_JP)) den += W.num2;
        if(!(fl & FL_JM)) den += W.num2;
        if(!skp) den += W.num3;
        if(!skm) den += W.num3;
    }
    const double src_max = den * cap;
    if(!isfinite(div_src)) div_src = I(0.0);
    else if(div_src > src_max) div_src = I(src_max);
    else if(div_src < -src_max) div_src = I(-src_max);
    I* src = ((i + j + k) & 1) ? src1 : src0;
    src[((size_t)i * jm + j) * hp + (k >> 1)] = div_src;
}

// Red-black split storage. The sweep domain's ring of km-1 meridians holds (km-1)/2 cells of each
// colour per (i, j) row; colour c = (i + j + k) & 1 lives in its own array X_c at compact index
// m = k >> 1 (row length hp, the seam cell k = km-1 at m = (km-1)/2). A colour pass then reads the
// other colour's array and its own source and flags with unit stride and writes its own array
// with whole cache lines. In a row where the own colour sits at k = 2m + off, the phi neighbours
// are the other colour's m - 1 + off and m + off (periodic on the ring); the i and j neighbours
// are the other colour's m.
template<typename T, typename U>
__global__ void k_rb_pack(const T* __restrict__ p, U* __restrict__ x0, U* __restrict__ x1,
        int jm, int km, int pitch, int hp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    U* x = ((i + j + k) & 1) ? x1 : x0;
    x[((size_t)i * jm + j) * hp + (k >> 1)] = (U)p[((size_t)i * jm + j) * pitch + k];
}
template<typename T, typename U>
__global__ void k_rb_unpack(T* __restrict__ p, const U* __restrict__ x0, const U* __restrict__ x1,
        int jm, int km, int pitch, int hp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const U* x = ((i + j + k) & 1) ? x1 : x0;
    p[((size_t)i * jm + j) * pitch + k] = (T)x[((size_t)i * jm + j) * hp + (k >> 1)];
}
__global__ void k_rb_flags(const uint8_t* __restrict__ flags, uint8_t* __restrict__ f0, uint8_t* __restrict__ f1,
        int jm, int km, int pitch, int hp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    uint8_t* f = ((i + j + k) & 1) ? f1 : f0;
    f[((size_t)i * jm + j) * hp + (k >> 1)] = flags[((size_t)i * jm + j) * pitch + k];
}

// One colour pass of a red-black sweep (PressureSolver.h:323-500) on the split storage. The domain
// boundary planes are the outer condition of the relaxation; the reference rewrites them after
// every sweep (PressureSolver.h:718-746). Here a plane value is only ever needed by its one
// interior neighbour, so from the second sweep on it is formed from the interior values it is
// defined by: a colour-0 reader forms it itself (both sources are unchanged at that point), and
// the colour-0 thread that owns the second source cell stores it just before updating that cell,
// for the colour-1 reader. The stored planes are therefore exactly what the reference would hold
// at every read; all planes are rewritten once at the end.
constexpr int RB_BM = 64, RB_BJ = 4, RB_NI = 1;

// A thread may own RB_NI consecutive levels of one (j, m) column of its colour, loading the other
// colour's column segment i0-1 .. i0+RB_NI once; measured on MI355X, one level per thread (the
// most waves in flight) is the fastest, 11.5 us per colour pass against 13.1 us with four.
template<typename T, typename S>
__global__ __launch_bounds__(RB_BM * RB_BJ) void k_rb_sweep(T* __restrict__ x0, T* __restrict__ x1,
        const S* __restrict__ s0, const S* __restrict__ s1,
        const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1, const PoisIJ<double>* __restrict__ pw,
        int im, int jm, int hp, int nm, int nmb, int njb, int nb, int colour, int ghost, double omega,
        double c43, double c13){
    const int b = xcd_remap(blockIdx.x, nb);
    const int mb = b % nmb, rest = b / nmb;
    const int jb = rest % njb, ib = rest / njb;
    const int m = mb * RB_BM + threadIdx.x;
    const int j = jb * RB_BJ + threadIdx.y + 1;
    const int i0 = ib * RB_NI + 1;
    if(m >= nm || j > jm - 2 || i0 > im - 2) return;
    T* __restrict__ own = colour ? x1 : x0;
    const T* __restrict__ oth = colour ? x0 : x1;
    const S* __restrict__ src = colour ? s1 : s0;
    const uint8_t* __restrict__ flg = colour ? f1 : f0;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const bool gw = ghost && colour == 0;      // ghost writes and on-the-fly ghost reads in this pass
    const int n_here = min(RB_NI, im - 2 - i0 + 1);
    T col[RB_NI + 2];
    #pragma unroll
    for(int t = 0; t < RB_NI + 2; t++){
        const int i = i0 - 1 + t;
        col[t] = (t <= n_here + 1) ? oth[((size_t)i * jm + j) * hp + m] : T(0);
    }
    #pragma unroll
    for(int t = 0; t < RB_NI; t++){
        if(t >= n_here) break;
        const int i = i0 + t;
        const size_t row = ((size_t)i * jm + j) * hp;
        const size_t q = row + m;
        const int off = (colour + i + j) & 1;
        const int mw = off ? m : (m == 0 ? nm - 1 : m - 1);
        const int me = off ? (m == nm - 1 ? 0 : m + 1) : m;
        const bool edge = i == 1 || i == 2 || i == im - 3 || i == im - 2 || j == 1 || j == 2 || j == jm - 3 || j == jm - 2;
        const bool need_pc = (omega != 1.0) || (gw && edge);
        const T pc = need_pc ? own[q] : T(0);
        if(gw){
            if(i == 2)      own[q - 2 * P] = (T)((T)c43 * col[t] - (T)c13 * pc);
            if(i == im - 3) own[q + 2 * P] = (T)((T)c43 * col[t + 2] - (T)c13 * pc);
            if(j == 2)      own[q - 2 * R] = (T)((T)c43 * oth[q - R] - (T)c13 * pc);
            if(j == jm - 3) own[q + 2 * R] = (T)((T)c43 * oth[q + R] - (T)c13 * pc);
        }
        const T pip = (i == im - 2 && gw) ? (T)((T)c43 * pc - (T)c13 * col[t]) : col[t + 2];
        const T pim = (i == 1 && gw)      ? (T)((T)c43 * pc - (T)c13 * col[t + 2]) : col[t];
        const T pjp = (j == jm - 2 && gw) ? (T)((T)c43 * pc - (T)c13 * oth[q - R]) : oth[q + R];
        const T pjm = (j == 1 && gw)      ? (T)((T)c43 * pc - (T)c13 * oth[q + R]) : oth[q - R];
        const T pkp = oth[row + me], pkm = oth[row + mw];
        const uint8_t fl = flg[q];
        const S sq = src[q];        // issued with the other loads, ahead of the flag tests
        const PoisIJ<double> W = pw[wave_uniform(i * jm + j)];     // a wavefront is one (i, j) row
        const T num1 = (T)W.num1, num2 = (T)W.num2, num3 = (T)W.num3;
        T acc, den;
        bool skip = false;
        if(!(fl & FL_SELF)){
            acc = den = T(0);
            if(!(fl & FL_IP)){ acc += num1 * pip; den += num1; }
            if(!(fl & FL_IM)){ acc += num1 * pim; den += num1; }
            if(!(fl & FL_JP)){ acc += num2 * pjp; den += num2; }
            if(!(fl & FL_JM)){ acc += num2 * pjm; den += num2; }
            if(!(fl & FL_KP)){ acc += num3 * pkp; den += num3; }
            if(!(fl & FL_KM)){ acc += num3 * pkm; den += num3; }
            skip = !(den > T(0));
        } else {
            acc = (pip + pim) * num1 + (pjp + pjm) * num2 + (pkp + pkm) * num3;
            den = (T)W.denom;
        }
        if(skip) continue;
        const T p_gs = (acc - (T)sq) / den;
        own[q] = (omega == 1.0) ? p_gs : (T)((T(1) - (T)omega) * pc + (T)omega * p_gs);
    }
}

// The colour pass with each thread owning two consecutive compact cells m, m+1 (m even) of one
// row: every neighbour row is read as one pair load, so a wavefront covers 128 cells with the
// vector-memory instructions the one-cell pass spends on 64. The arithmetic of a cell is that
// pass's. Needs an even ring (nm).
template<typename T, typename S>
__device__ __forceinline__ void rb_cell(T* __restrict__ own, const S sq,
        const uint8_t fl, const PoisIJ<double>& W, size_t q, size_t P, size_t R, int i, int j, int im, int jm,
        bool gw, double omega, double c43, double c13, T oim, T oip, T ojm, T ojp, T pkp, T pkm){
    const bool edge = i == 1 || i == 2 || i == im - 3 || i == im - 2 || j == 1 || j == 2 || j == jm - 3 || j == jm - 2;
    const bool need_pc = (omega != 1.0) || (gw && edge);
    const T pc = need_pc ? own[q] : T(0);
    if(gw){
        if(i == 2)      own[q - 2 * P] = (T)((T)c43 * oim - (T)c13 * pc);
        if(i == im - 3) own[q + 2 * P] = (T)((T)c43 * oip - (T)c13 * pc);
        if(j == 2)      own[q - 2 * R] = (T)((T)c43 * ojm - (T)c13 * pc);
        if(j == jm - 3) own[q + 2 * R] = (T)((T)c43 * ojp - (T)c13 * pc);
    }
    const T pip = (i == im - 2 && gw) ? (T)((T)c43 * pc - (T)c13 * oim) : oip;
    const T pim = (i == 1 && gw)      ? (T)((T)c43 * pc - (T)c13 * oip) : oim;
    const T pjp = (j == jm - 2 && gw) ? (T)((T)c43 * pc - (T)c13 * ojm) : ojp;
    const T pjm = (j == 1 && gw)      ? (T)((T)c43 * pc - (T)c13 * ojp) : ojm;
    const T num1 = (T)W.num1, num2 = (T)W.num2, num3 = (T)W.num3;
    T acc, den;
    if(!(fl & FL_SELF)){
        acc = den = T(0);
        if(!(fl & FL_IP)){ acc += num1 * pip; den += num1; }
        if(!(fl & FL_IM)){ acc += num1 * pim; den += num1; }
        if(!(fl & FL_JP)){ acc += num2 * pjp; den += num2; }
        if(!(fl & FL_JM)){ acc += num2 * pjm; den += num2; }
        if(!(fl & FL_KP)){ acc += num3 * pkp; den += num3; }
        if(!(fl & FL_KM)){ acc += num3 * pkm; den += num3; }
        if(!(den > T(0))) return;
    } else {
        acc = (pip + pim) * num1 + (pjp + pjm) * num2 + (pkp + pkm) * num3;
        den = (T)W.denom;
    }
    const T p_gs = (acc - (T)sq) / den;
    own[q] = (omega == 1.0) ? p_gs : (T)((T(1) - (T)omega) * pc + (T)omega * p_gs);
}

// rb_cell where the pass writes no ghost plane at this cell and reads none on the fly (the pass
// is not a ghost pass, or the cell is at least three rows from every outer plane): the neighbour
// values are the stored ones. A fluid cell with six fluid neighbours sums the same six terms in the
// same order from 0 and takes the row's precomputed denominator of that sum.
template<typename T, typename S>
__device__ __forceinline__ void rb_cell_plain(T* __restrict__ own, const S sq, const uint8_t fl,
        const PoisIJ<double>& W, size_t q, double omega, T pim, T pip, T pjm, T pjp, T pkp, T pkm){
    const T num1 = (T)W.num1, num2 = (T)W.num2, num3 = (T)W.num3;
    T acc, den;
    if((fl & 0x7f) == 0){
        acc = T(0) + num1 * pip;
        acc += num1 * pim;
        acc += num2 * pjp;
        acc += num2 * pjm;
        acc += num3 * pkp;
        acc += num3 * pkm;
        den = std::is_same<T, float>::value ? (T)W.den_fluid32 : (T)W.den_fluid;
    } else if(!(fl & FL_SELF)){
        acc = den = T(0);
        if(!(fl & FL_IP)){ acc += num1 * pip; den += num1; }
        if(!(fl & FL_IM)){ acc += num1 * pim; den += num1; }
        if(!(fl & FL_JP)){ acc += num2 * pjp; den += num2; }
        if(!(fl & FL_JM)){ acc += num2 * pjm; den += num2; }
        if(!(fl & FL_KP)){ acc += num3 * pkp; den += num3; }
        if(!(fl & FL_KM)){ acc += num3 * pkm; den += num3; }
        if(!(den > T(0))) return;
    } else {
        acc = (pip + pim) * num1 + (pjp + pjm) * num2 + (pkp + pkm) * num3;
        den = (T)W.denom;
    }
    const T p_gs = (acc - (T)sq) / den;
    own[q] = (omega == 1.0) ? p_gs : (T)((T(1) - (T)omega) * own[q] + (T)omega * p_gs);
}

// t / d and t % d for 0 <= t < 2^22 and 1 <= d < 2^22 through a float reciprocal of d: the
// quotient it gives is off by at most one, which the remainder's range corrects.
__device__ __forceinline__ int div_small(int t, int d, float inv_d, int& rem){
    int qt = (int)((float)t * inv_d);
    int r = t - qt * d;
    if(r < 0){ qt -= 1; r += d; }
    else if(r >= d){ qt += 1; r -= d; }
    rem = r;
    return qt;
}

// Does a colour pass write or read ghost planes at row (i, j)? Only the ghost pass (colour 0 of a
// sweep after the first) does, and only within two rows of an outer plane.
__device__ __forceinline__ bool rb_ghost_row(bool gw, int i, int j, int im, int jm){
    return gw && (i <= 2 || i >= im - 3 || j <= 2 || j >= jm - 3);
}

template<typename T, typename S>
__global__ __launch_bounds__(RB_BM * RB_BJ) void k_rb_sweep2(T* __restrict__ x0, T* __restrict__ x1,
        const S* __restrict__ s0, const S* __restrict__ s1,
        const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1, const PoisIJ<double>* __restrict__ pw,
        int im, int jm, int hp, int nm, int nmb, int njb, int nb, int colour, int ghost, double omega,
        double c43, double c13){
    using V = typename Pair<T>::V;
    const int b = xcd_remap(blockIdx.x, nb);
    const int mb = b % nmb, rest = b / nmb;
    const int jb = rest % njb, ib = rest / njb;
    const int m = 2 * (mb * RB_BM + threadIdx.x);
    const int j = jb * RB_BJ + threadIdx.y + 1;
    const int i = ib + 1;
    if(m >= nm || j > jm - 2) return;
    T* __restrict__ own = colour ? x1 : x0;
    const T* __restrict__ oth = colour ? x0 : x1;
    const S* __restrict__ src = colour ? s1 : s0;
    const uint8_t* __restrict__ flg = colour ? f1 : f0;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp;
    const size_t q = row + m;
    const bool gw = ghost && colour == 0;
    const V vim = *reinterpret_cast<const V*>(oth + q - P), vip = *reinterpret_cast<const V*>(oth + q + P);
    const V vjm = *reinterpret_cast<const V*>(oth + q - R), vjp = *reinterpret_cast<const V*>(oth + q + R);
    const V vc = *reinterpret_cast<const V*>(oth + q);               // the other colour at m, m+1
    // the own colour's sources, loaded with the neighbours: every path reads them, and issuing them
    // here saves the second round trip to memory a load after the flag tests would cost
    const typename Pair<S>::V sv = *reinterpret_cast<const typename Pair<S>::V*>(src + q);
    const int off = (colour + i + j) & 1;
    // the phi neighbours: off = 1: m | m+1 for cell m, m+1 | m+2 for cell m+1; off = 0: m-1 | m, m | m+1
    const T edge = off ? oth[row + (m + 2 == nm ? 0 : m + 2)] : oth[row + (m == 0 ? nm - 1 : m - 1)];
    const uint16_t fl2 = *reinterpret_cast<const uint16_t*>(flg + q);
    const PoisIJ<double> W = pw[wave_uniform(i * jm + j)];
    T pkm0, pkp0, pkm1, pkp1;
    if(off){ pkm0 = vc.x; pkp0 = vc.y; pkm1 = vc.y; pkp1 = edge; }
    else   { pkm0 = edge; pkp0 = vc.x; pkm1 = vc.x; pkp1 = vc.y; }
    if(rb_ghost_row(gw, i, j, im, jm)){
        rb_cell<T, S>(own, sv.x, (uint8_t)(fl2 & 0xff), W, q, P, R, i, j, im, jm, gw, omega, c43, c13,
                      vim.x, vip.x, vjm.x, vjp.x, pkp0, pkm0);
        rb_cell<T, S>(own, sv.y, (uint8_t)(fl2 >> 8), W, q + 1, P, R, i, j, im, jm, gw, omega, c43, c13,
                      vim.y, vip.y, vjm.y, vjp.y, pkp1, pkm1);
    } else {
        rb_cell_plain<T, S>(own, sv.x, (uint8_t)(fl2 & 0xff), W, q, omega, vim.x, vip.x, vjm.x, vjp.x, pkp0, pkm0);
        rb_cell_plain<T, S>(own, sv.y, (uint8_t)(fl2 >> 8), W, q + 1, omega, vim.y, vip.y, vjm.y, vjp.y, pkp1, pkm1);
    }
}

// The multi-cell pass with the (row, group) index flattened over the lanes, so that no lane idles
// at the end of a row: each thread owns N consecutive compact cells m..m+N-1 (m a multiple of N),
// read as N/2 pair loads per neighbour row; a wavefront may span two rows and reads the row's
// weights per lane. Needs nm % N == 0.
template<typename T, typename S, int N>
__global__ __launch_bounds__(256) void k_rb_sweep2f(T* __restrict__ x0, T* __restrict__ x1,
        const S* __restrict__ s0, const S* __restrict__ s1,
        const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1, const PoisIJ<double>* __restrict__ pw,
        int im, int jm, int hp, int nm, int ntb, int nb, int colour, int ghost, double omega,
        double c43, double c13, float inv_ngrp){
    using V = typename Pair<T>::V;
    static_assert(N == 2 || N == 4, "two or four cells per thread");
    const int b = xcd_remap(blockIdx.x, nb);
    // levels fastest: a group's neighbours at i +- 1 are read by the workgroups dispatched beside it
    const int i = b % (im - 2) + 1, tb = b / (im - 2);
    const int ngrp = nm / N;
    const int t = tb * 256 + threadIdx.x;
    int mg;
    const int j = div_small(t, ngrp, inv_ngrp, mg) + 1;
    const int m = N * mg;
    if(j > jm - 2) return;
    T* __restrict__ own = colour ? x1 : x0;
    const T* __restrict__ oth = colour ? x0 : x1;
    const S* __restrict__ src = colour ? s1 : s0;
    const uint8_t* __restrict__ flg = colour ? f1 : f0;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp;
    const size_t q = row + m;
    const bool gw = ghost && colour == 0;
    T vim[N], vip[N], vjm[N], vjp[N], vc[N];
    S sv[N];        // the own colour's sources, loaded with the neighbours (one round trip to memory)
    #pragma unroll
    for(int h = 0; h < N; h += 2){
        const V a = *reinterpret_cast<const V*>(oth + q + h - P), bq = *reinterpret_cast<const V*>(oth + q + h + P);
        const V c = *reinterpret_cast<const V*>(oth + q + h - R), d = *reinterpret_cast<const V*>(oth + q + h + R);
        const V e = *reinterpret_cast<const V*>(oth + q + h);
        const typename Pair<S>::V f = *reinterpret_cast<const typename Pair<S>::V*>(src + q + h);
        vim[h] = a.x; vim[h + 1] = a.y; vip[h] = bq.x; vip[h + 1] = bq.y;
        vjm[h] = c.x; vjm[h + 1] = c.y; vjp[h] = d.x; vjp[h + 1] = d.y;
        vc[h] = e.x; vc[h + 1] = e.y;
        sv[h] = f.x; sv[h + 1] = f.y;
    }
    const int off = (colour + i + j) & 1;
    // the phi neighbours: off = 1: cell m+r reads m+r (k-1) and m+r+1 (k+1); off = 0: m+r-1, m+r
    const T edge = off ? oth[row + (m + N == nm ? 0 : m + N)] : oth[row + (m == 0 ? nm - 1 : m - 1)];
    uint32_t flw;
    if constexpr(N == 2) flw = *reinterpret_cast<const uint16_t*>(flg + q);
    else flw = *reinterpret_cast<const uint32_t*>(flg + q);
    const PoisIJ<double> W = pw[(size_t)i * jm + j];
    const bool ghost_row = rb_ghost_row(gw, i, j, im, jm);
    constexpr uint32_t all_fluid_mask = (N == 2) ? 0x7f7fu : 0x7f7f7f7fu;
    if(!ghost_row && (flw & all_fluid_mask) == 0 && omega == 1.0){
        // every cell of the group fluid with fluid neighbours: the sources read and the results
        // written as pairs
        const T num1 = (T)W.num1, num2 = (T)W.num2, num3 = (T)W.num3;
        const T den = std::is_same<T, float>::value ? (T)W.den_fluid32 : (T)W.den_fluid;
        #pragma unroll
        for(int h = 0; h < N; h += 2){
            T res[2];
            #pragma unroll
            for(int e = 0; e < 2; e++){
                const int r = h + e;
                T pkm, pkp;
                if(off){ pkm = vc[r]; pkp = (r + 1 < N) ? vc[r + 1] : edge; }
                else   { pkm = (r > 0) ? vc[r - 1] : edge; pkp = vc[r]; }
                T acc = T(0) + num1 * vip[r];
                acc += num1 * vim[r];
                acc += num2 * vjp[r];
                acc += num2 * vjm[r];
                acc += num3 * pkp;
                acc += num3 * pkm;
                res[e] = (acc - (T)sv[r]) / den;
            }
            V o; o.x = res[0]; o.y = res[1];
            *reinterpret_cast<V*>(own + q + h) = o;
        }
        return;
    }
    #pragma unroll
    for(int r = 0; r < N; r++){
        T pkm, pkp;
        if(off){ pkm = vc[r]; pkp = (r + 1 < N) ? vc[r + 1] : edge; }
        else   { pkm = (r > 0) ? vc[r - 1] : edge; pkp = vc[r]; }
        const uint8_t fl = (uint8_t)((flw >> (8 * r)) & 0xff);
        if(ghost_row)
            rb_cell<T, S>(own, sv[r], fl, W, q + r, P, R, i, j, im, jm, gw, omega, c43, c13,
                          vim[r], vip[r], vjm[r], vjp[r], pkp, pkm);
        else
            rb_cell_plain<T, S>(own, sv[r], fl, W, q + r, omega, vim[r], vip[r], vjm[r], vjp[r], pkp, pkm);
    }
}

// ---- the FP32-refined solve: FP32 sweeps on the correction inside FP64 refinement ----
// A red-black sweep is an affine map, so from a start p0 the iterates are p0 + d with d relaxed by
// the same sweep on zero start and source r0 = src + den p0 - sum(w p0_nb), the cell's FP64
// Gauss-Seidel residual of p0 (solid cells: all six neighbours with the full denominator, as in
// the sweep). Each refinement round forms r0 in FP64, relaxes d in FP32 and adds it in FP64; with
// the outer planes of p0 consistent (c43 p1 - c13 p2) the planes of d start consistent at zero,
// so the ghost-plane scheme applies from the round's first sweep.
__global__ __launch_bounds__(256) void k_rb_residual(const double* __restrict__ x0, const double* __restrict__ x1,
        const double* __restrict__ s0, const double* __restrict__ s1,
        const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1, const PoisIJ<double>* __restrict__ pw,
        float* __restrict__ r0, float* __restrict__ r1, int im, int jm, int hp, int nm){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    const int i = (blockIdx.z >> 1) + 1, colour = blockIdx.z & 1;
    if(m >= nm || j > jm - 2 || i > im - 2) return;
    const double* own = colour ? x1 : x0;
    const double* oth = colour ? x0 : x1;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp;
    const size_t q = row + m;
    const int off = (colour + i + j) & 1;
    const int mw = off ? m : (m == 0 ? nm - 1 : m - 1);
    const int me = off ? (m == nm - 1 ? 0 : m + 1) : m;
    const double pip = oth[q + P], pim = oth[q - P], pjp = oth[q + R], pjm = oth[q - R];
    const double pkp = oth[row + me], pkm = oth[row + mw];
    const uint8_t fl = (colour ? f1 : f0)[q];
    const PoisIJ<double> W = pw[(size_t)i * jm + j];
    double acc, den;
    if(!(fl & FL_SELF)){
        acc = den = 0.0;
        if(!(fl & FL_IP)){ acc += W.num1 * pip; den += W.num1; }
        if(!(fl & FL_IM)){ acc += W.num1 * pim; den += W.num1; }
        if(!(fl & FL_JP)){ acc += W.num2 * pjp; den += W.num2; }
        if(!(fl & FL_JM)){ acc += W.num2 * pjm; den += W.num2; }
        if(!(fl & FL_KP)){ acc += W.num3 * pkp; den += W.num3; }
        if(!(fl & FL_KM)){ acc += W.num3 * pkm; den += W.num3; }
    } else {
        acc = (pip + pim) * W.num1 + (pjp + pjm) * W.num2 + (pkp + pkm) * W.num3;
        den = W.denom;
    }
    (colour ? r1 : r0)[q] = (float)((colour ? s1 : s0)[q] + den * own[q] - acc);
}

// p0 += d over the sweep domain, both colours; d is cleared for the next round
__global__ __launch_bounds__(256) void k_rb_add(double* __restrict__ x0, double* __restrict__ x1,
        float* __restrict__ d0, float* __restrict__ d1, int im, int jm, int hp, int nm){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1;
    const int i = (blockIdx.z >> 1) + 1, colour = blockIdx.z & 1;
    if(m >= nm || j > jm - 2 || i > im - 2) return;
    const size_t q = ((size_t)i * jm + j) * hp + m;
    float* d = colour ? d1 : d0;
    (colour ? x1 : x0)[q] += (double)d[q];
    d[q] = 0.0f;
}

__global__ void k_fill_float(float* __restrict__ x, size_t n, float v){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) x[q] = v;
}

// the outer planes of the split iterate, i planes then j planes (PressureSolver.h:718-746)
__global__ void k_rb_planes_i(double* __restrict__ x0, double* __restrict__ x1, int im, int jm, int hp, int nm,
        double c43, double c13){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, side = blockIdx.z;
    if(m >= nm) return;
    const size_t P = (size_t)jm * hp;
    // plane cell (ia, j, k) has the colour of (ib2, j, k); (ib1, j, k) the other one
    const int ia = side ? im - 1 : 0, ib1 = side ? im - 2 : 1, ib2 = side ? im - 3 : 2;
    for(int c = 0; c < 2; c++){
        double* same = c ? x1 : x0;
        const double* other = c ? x0 : x1;
        const size_t base = (size_t)j * hp + m;
        same[base + ia * P] = c43 * other[base + ib1 * P] - c13 * same[base + ib2 * P];
    }
}
__global__ void k_rb_planes_j(double* __restrict__ x0, double* __restrict__ x1, int im, int jm, int hp, int nm,
        double c43, double c13){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, side = blockIdx.z;
    if(m >= nm) return;
    const size_t R = (size_t)hp;
    const int ja = side ? jm - 1 : 0, jb1 = side ? jm - 2 : 1, jb2 = side ? jm - 3 : 2;
    for(int c = 0; c < 2; c++){
        double* same = c ? x1 : x0;
        const double* other = c ? x0 : x1;
        const size_t base = (size_t)i * jm * hp + m;
        same[base + ja * R] = c43 * other[base + jb1 * R] - c13 * same[base + jb2 * R];
    }
}

// the outer boundary after the last sweep: i planes, then j planes, then the seam copy
template<typename T>
__global__ void k_poisson_planes_i(T* __restrict__ p, int im, int jm, int km, int pitch, size_t plane,
        double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q = (size_t)j * pitch + k;
    p[q] = (T)c43 * p[q + plane] - (T)c13 * p[q + 2 * plane];
    p[q + (size_t)(im - 1) * plane] = (T)c43 * p[q + (size_t)(im - 2) * plane] - (T)c13 * p[q + (size_t)(im - 3) * plane];
}
template<typename T>
__global__ void k_poisson_planes_j(T* __restrict__ p, int im, int jm, int km, int pitch, size_t plane,
        double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(k >= km) return;
    const size_t q = (size_t)i * plane + k, R = (size_t)pitch;
    p[q] = (T)c43 * p[q + R] - (T)c13 * p[q + 2 * R];
    p[q + (size_t)(jm - 1) * R] = (T)c43 * p[q + (size_t)(jm - 2) * R] - (T)c13 * p[q + (size_t)(jm - 3) * R];
}
template<typename T>
__global__ void k_seam_copy(T* __restrict__ p, int im, int jm, int km, int pitch){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(j >= jm) return;
    const size_t q = ((size_t)i * jm + j) * pitch;
    p[q + km - 1] = p[q];
}

// ---------------------------------------------------------------------------------------------
// Shapiro filter passes (lib/Utils.cpp:221-453): each pass reads a snapshot and writes a second
// buffer; a cell outside the fluid (below i_topography) keeps its value.
// ---------------------------------------------------------------------------------------------
template<typename T>
__global__ __launch_bounds__(256) void k_shapiro_k(const T* __restrict__ in, T* __restrict__ out,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, int order, double strength){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t r0 = ((size_t)i * jm + j) * pitch;
    const int kk = (k == km - 1) ? 0 : k;                  // the seam copy takes the value of k = 0
    const int* it = itopo + (size_t)j * km;
    auto fl = [&](int kx){ return i >= max(it[kx], 0); };
    const T c = in[r0 + kk];
    T res = c;
    if(fl(kk)){
        const int n_k = km - 1;
        const int km1 = (kk == 0) ? km - 2 : kk - 1, kp1 = (kk == km - 2) ? 0 : kk + 1;
        const T coeff = (T)(strength * 0.25);
        if(order == 2){
            const T a = fl(km1) ? in[r0 + km1] : c;
            const T b = fl(kp1) ? in[r0 + kp1] : c;
            res = c + coeff * (a - T(2) * c + b);
        } else if(fl(km1) && fl(kp1)){
            const int km2 = (kk - 2 + n_k) % n_k, kp2 = (kk + 2) % n_k;
            if(fl(km2) && fl(kp2)){
                const T d4 = in[r0 + kp2] - T(4) * in[r0 + kp1] + T(6) * c - T(4) * in[r0 + km1] + in[r0 + km2];
                res = c - (T)(strength / 16.0) * d4;
            } else {
                res = c + coeff * (in[r0 + km1] - T(2) * c + in[r0 + kp1]);
            }
        }
    }
    out[r0 + k] = res;
}

template<typename T>
__global__ __launch_bounds__(256) void k_shapiro_j(const T* __restrict__ in, T* __restrict__ out,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, int order, double strength){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t R = (size_t)pitch;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    auto fl = [&](int jj){ return i >= max(itopo[(size_t)jj * km + k], 0); };
    const T c = in[q];
    T res = c;
    if(fl(j)){
        const T coeff = (T)(strength * 0.25);
        if(order == 2){
            const int jm1 = max(j - 1, 0), jp1 = min(j + 1, jm - 1);
            const T a = fl(jm1) ? in[q - (size_t)(j - jm1) * R] : c;
            const T b = fl(jp1) ? in[q + (size_t)(jp1 - j) * R] : c;
            res = c + coeff * (a - T(2) * c + b);
        } else {
            const bool fm1 = (j - 1 >= 0) && fl(j - 1), fp1 = (j + 1 <= jm - 1) && fl(j + 1);
            if(fm1 && fp1){
                const bool fm2 = (j - 2 >= 0) && fl(j - 2), fp2 = (j + 2 <= jm - 1) && fl(j + 2);
                if(fm2 && fp2){
                    const T d4 = in[q + 2 * R] - T(4) * in[q + R] + T(6) * c - T(4) * in[q - R] + in[q - 2 * R];
                    res = c - (T)(strength / 16.0) * d4;
                } else {
                    res = c + coeff * (in[q - R] - T(2) * c + in[q + R]);
                }
            }
        }
    }
    out[q] = res;
}

template<typename T>
__global__ __launch_bounds__(256) void k_shapiro_i(const T* __restrict__ in, T* __restrict__ out,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, size_t plane, int order, double strength){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const int i_surf = max(itopo[(size_t)j * km + k], 0);
    const T c = in[q];
    T res = c;
    if(i >= i_surf){
        const T coeff = (T)(strength * 0.25);
        if(order == 2){
            const int im1 = max(i - 1, i_surf), ip1 = min(i + 1, im - 1);
            res = c + coeff * (in[q - (size_t)(i - im1) * plane] - T(2) * c + in[q + (size_t)(ip1 - i) * plane]);
        } else {
            const bool fm1 = (i - 1 >= i_surf), fp1 = (i + 1 <= im - 1);
            if(fm1 && fp1){
                const bool fm2 = (i - 2 >= i_surf), fp2 = (i + 2 <= im - 1);
                if(fm2 && fp2){
                    const T d4 = in[q + 2 * plane] - T(4) * in[q + plane] + T(6) * c - T(4) * in[q - plane] + in[q - 2 * plane];
                    res = c - (T)(strength / 16.0) * d4;
                } else {
                    res = c + coeff * (in[q - plane] - T(2) * c + in[q + plane]);
                }
            }
        }
    }
    out[q] = res;
}

// The order-2 pass of all three directions in one kernel: the phi pass, then the theta pass on its
// result, then the radial pass on that, each cell's arithmetic the one of k_shapiro_k, _j, _i in
// sequence (so the result is bit-identical to the three launches). A thread owns SH_L levels
// [i0, i0 + SH_L) of a (j, k) column: it forms the phi-filtered values of rows j-1, j, j+1 at those
// levels and the two beside them, the theta-filtered values of row j from those, and the radial
// results from three consecutive theta-filtered values. The workgroup first stages the raw tile it
// reads (SH_L + 2 levels, SH_BJ + 2 rows, SH_BK + 2 columns, each clamped or wrapped as the cells
// read it) in LDS with coalesced loads, about nine per thread instead of 54 scattered ones. The
// field is read once and written once instead of three times each.
constexpr int SH_BK = 64, SH_BJ = 4;
#ifndef SH_L
#define SH_L 4                                      // levels per thread
#endif
constexpr int SH_TL = SH_L + 2, SH_TJ = SH_BJ + 2, SH_TK = SH_BK + 2;

template<typename T>
__global__ __launch_bounds__(SH_BK * SH_BJ) void k_shapiro2_fused(const T* __restrict__ in, T* __restrict__ out,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, size_t plane, int nkb, int njb, int nch,
        T coeff){
    __shared__ T tile[SH_TL][SH_TJ][SH_TK];
    const int nb = nkb * njb * nch;
    // level ranges fastest, each XCD a contiguous range of blocks: the workgroups that share a
    // tile's outer levels and rows run beside each other on one L2
    const int b = xcd_remap(blockIdx.x, nb);
    const int ch = b % nch, rest = b / nch;
    const int kb = rest % nkb, jb = rest / nkb;
    const int k0 = kb * SH_BK, j0 = jb * SH_BJ, i0 = ch * SH_L;
    const int n = km - 1;                                  // distinct columns of the ring
    // tile slot (l, r, s): level clamp(i0 - 1 + l), row clamp(j0 - 1 + r), column wrap(k0 - 1 + s)
    // over the ring (the seam k = km - 1 reads column 0 and its neighbours km - 2 and 1, as the
    // column km - 1 + 1 = km wraps to 1)
    // the surface levels first (the masks of the cells the thread filters), so that their loads and
    // the tile's are in flight together
    const int k = k0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool own = k < km && j < jm;
    const int kc = own ? k : 0, jc = own ? j : 0;
    const int kk = (kc == km - 1) ? 0 : kc;                // the seam copy takes the value of k = 0
    const int km1 = (kk == 0) ? km - 2 : kk - 1, kp1 = (kk == km - 2) ? 0 : kk + 1;
    const int jm1 = max(jc - 1, 0), jp1 = min(jc + 1, jm - 1);
    auto surf = [&](int jj, int kx){ return max(itopo[(size_t)jj * km + kx], 0); };
    const int jr[3] = { jm1, jc, jp1 };
    int sc[3], sm[3], sp[3];
    #pragma unroll
    for(int r = 0; r < 3; r++){ sc[r] = surf(jr[r], kk); sm[r] = surf(jr[r], km1); sp[r] = surf(jr[r], kp1); }
    const int tm = surf(jm1, kc), tc = surf(jc, kc), tp = surf(jp1, kc);
    const int tid = threadIdx.y * SH_BK + threadIdx.x;
    for(int e = tid; e < SH_TL * SH_TJ * SH_TK; e += SH_BK * SH_BJ){
        const int s = e % SH_TK, rr = (e / SH_TK) % SH_TJ, l = e / (SH_TK * SH_TJ);
        const int i = min(max(i0 - 1 + l, 0), im - 1);
        const int jj = min(max(j0 - 1 + rr, 0), jm - 1);
        const int kx = ((k0 - 1 + s) % n + n) % n;
        tile[l][rr][s] = in[((size_t)i * jm + jj) * pitch + kx];
    }
    __syncthreads();
    if(!own) return;
    const T two = T(2);
    // the thread's column sits at slot threadIdx.x + 1 (its kk), with km1 and kp1 beside it; its
    // rows j - 1, j, j + 1 at slots threadIdx.y .. threadIdx.y + 2
    const int sk = threadIdx.x + 1, sj = threadIdx.y;
    // theta-filtered values of row j at levels i0 - 1 .. i0 + SH_L (clamped into the column; the
    // clamped ones are never used)
    T th[SH_L + 2];
    #pragma unroll
    for(int l = 0; l < SH_L + 2; l++){
        const int i = min(max(i0 - 1 + l, 0), im - 1);
        T ph[3];
        // the filtered values are formed unconditionally and selected
        #pragma unroll
        for(int r = 0; r < 3; r++){                        // phi pass of row jr[r] (k_shapiro_k)
            const T c = tile[l][sj + r][sk], am = tile[l][sj + r][sk - 1], ap = tile[l][sj + r][sk + 1];
            const T a = (i >= sm[r]) ? am : c;
            const T bb = (i >= sp[r]) ? ap : c;
            const T f = c + coeff * (a - two * c + bb);
            ph[r] = (i < sc[r]) ? c : f;
        }
        const T c = ph[1];                                 // theta pass (k_shapiro_j)
        const T a = (i >= tm) ? ph[0] : c;
        const T bb = (i >= tp) ? ph[2] : c;
        const T f = c + coeff * (a - two * c + bb);
        th[l] = (i < tc) ? c : f;
    }
    // radial pass (k_shapiro_i, order 2)
    const int i_surf = tc;
    const size_t col = (size_t)j * pitch + k;
    #pragma unroll
    for(int l = 1; l <= SH_L; l++){
        const int i = i0 - 1 + l;
        if(i >= im) break;
        const T cur = th[l];
        const T lo = (i - 1 >= i_surf) ? th[l - 1] : cur;
        const T hi = (i + 1 <= im - 1) ? th[l + 1] : cur;
        const T f = cur + coeff * (lo - two * cur + hi);
        out[col + (size_t)i * plane] = (i >= i_surf) ? f : cur;
    }
}

// ---------------------------------------------------------------------------------------------
// computeMixtureDensity (InitValues_Jup.cpp:717-740), every cell
// ---------------------------------------------------------------------------------------------
template<typename T>
__global__ __launch_bounds__(256) void k_mixture_density(const T* __restrict__ t, const T* __restrict__ ps,
        const T* __restrict__ pd, T* __restrict__ rho, size_t n, double t_ref, double pdyn_w, double ptb, double R_mix){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    using A = acc_t<T>;
    const A T_ = (A)t[q] * t_ref;
    if(!(T_ > 0.0)){ rho[q] = T(0); return; }
    const A p_pa = ((A)ps[q] + pdyn_w * (A)pd[q] * ptb) * 1.0e5;
    const A r = p_pa / (R_mix * T_);
    rho[q] = isfinite(r) ? (T)r : T(0);
}

// ---------------------------------------------------------------------------------------------
// Boundary conditions after the RK4 step
// ---------------------------------------------------------------------------------------------
constexpr int MAXBC = 40;           // the longest list (phi) holds 32 fields in the moist profile
template<typename T>
struct BCLists {
    T* radius[MAXBC]; int n_radius;                 // bc_fields_radius (BC_Jup.h:272-290)
    T* theta_ex[MAXBC]; int n_theta_ex;             // bc_fields_theta_extrap (:293-312)
    T* theta_zero[MAXBC]; int n_theta_zero;         // bc_fields_theta_zero (:318-325)
    T* phi[MAXBC]; int n_phi;                       // bc_fields_phi (:328-350)
    T* species[MAXBC]; int n_species;               // the wall copy / positivity set (BC_Jup.h:232-262)
    T* solid[MAXBC]; int n_solid;                   // bcSolidGround scalars (BC_Jup.h:469-487)
    T *u, *v, *w, *t;
    // what bcRadius does to a radius-list field after its Neumann planes: 0 nothing, 1 the rigid
    // lid (u), 2 / 3 the top sponge (v / w), 4 the species wall copy and positivity
    uint8_t radius_role[MAXBC];
};

// the zonal means of v, w at i = im-3, im-2 before the top sponge, k = 0..km-2 in order: one
// wavefront per (layer, field, j) row stages the row in LDS with coalesced loads, then one lane
// adds it up in the reference's serial order
constexpr int ZM_MAXK = 1024;
template<typename T>
__global__ __launch_bounds__(64) void k_bc_zonal_means(const T* __restrict__ v, const T* __restrict__ w,
        acc_t<T>* __restrict__ zm, int im, int jm, int km, int pitch){
    __shared__ acc_t<T> row[ZM_MAXK];
    const int j = blockIdx.x;
    const int sel = blockIdx.y;                    // 0..3 = (layer 0 v, layer 0 w, layer 1 v, layer 1 w)
    const int l = sel >> 1;
    const T* X = (sel & 1) ? w : v;
    const size_t r0 = ((size_t)(im - 3 + l) * jm + j) * pitch;
    for(int k = threadIdx.x; k < km - 1; k += 64) row[k] = (acc_t<T>)X[r0 + k];
    __syncthreads();
    if(threadIdx.x != 0) return;
    acc_t<T> s = 0.0;
    for(int k = 0; k < km - 1; k++) s += row[k];
    zm[(size_t)sel * jm + j] = s / (double)(km - 1);
}

// BC_Jup::bcRadius (BC_Jup.h:185-263) with the shared bcRadius (BoundaryConditions.h:151-286),
// ATJUP defaults: 2-point Neumann, rigid lid on u, TOP_TAPER 3, species wall copy, positivity.
template<typename T>
__global__ __launch_bounds__(256) void k_bc_radius(const BCLists<T> L, const acc_t<T>* __restrict__ zm,
        int im, int jm, int km, int pitch, size_t plane, double c43, double c13,
        float* __restrict__ lo_v, float* __restrict__ lo_w){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, f = blockIdx.z;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    const size_t qt = q0 + (size_t)(im - 1) * plane;
    T* X = L.radius[f];
    X[q0] = (T)c43 * X[q0 + plane] - (T)c13 * X[q0 + 2 * plane];
    X[qt] = (T)c43 * X[qt - plane] - (T)c13 * X[qt - 2 * plane];
    const int role = L.radius_role[f];
    if(role == 1){                                 // rigid lid
        X[q0] = T(0);
        X[qt] = T(0);
    } else if(role == 2 || role == 3){             // top sponge, zonal-mean preserving
        const acc_t<T> m3 = zm[(size_t)(role - 2) * jm + j], m2 = zm[(size_t)(role) * jm + j];
        X[qt] = T(0);
        X[qt - plane] *= (T)(1.0 / 3.0);
        X[qt - 2 * plane] *= (T)(2.0 / 3.0);
        X[qt - 2 * plane] += (T)((1.0 / 3.0) * m3);
        X[qt - plane] += (T)((2.0 / 3.0) * m2);
        X[qt] = (T)m2;
        float* lo = (role == 2) ? lo_v : lo_w;
        if(lo){                                    // the sponge's levels carry no low words
            lo[qt - plane] = 0.0f;
            lo[qt - 2 * plane] = 0.0f;
        }
    } else if(role == 4){                          // species: wall copy, positivity
        X[q0] = X[q0 + plane];
        if(X[q0] < T(0)) X[q0] = T(0);
        if(X[qt] < T(0)) X[qt] = T(0);
    }
}

// bcTheta with POLE_FILL 1 (two open rows) and POLE_COPY 1 (BoundaryConditions.h:292-391)
template<typename T>
__global__ __launch_bounds__(256) void k_bc_theta(const BCLists<T> L, int im, int jm, int km, int pitch){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, f = blockIdx.z;
    if(k >= km) return;
    const size_t R = (size_t)pitch;
    const size_t qs = (size_t)i * jm * pitch + k;                 // j = 0
    const size_t qn = qs + (size_t)(jm - 1) * R;                  // j = jm-1
    if(f < L.n_theta_ex){
        T* X = L.theta_ex[f];
        const T a = X[qs + 3 * R], b = X[qn - 3 * R];
        X[qs + R] = a; X[qs + 2 * R] = a; X[qs] = a;
        X[qn - R] = b; X[qn - 2 * R] = b; X[qn] = b;
    } else {
        T* X = L.theta_zero[f - L.n_theta_ex];
        const T a = X[qs + 3 * R], b = X[qn - 3 * R];
        X[qs + R] = (T)(1.0 / 3.0) * a;
        X[qs + 2 * R] = (T)(2.0 / 3.0) * a;
        X[qn - R] = (T)(1.0 / 3.0) * b;
        X[qn - 2 * R] = (T)(2.0 / 3.0) * b;
        X[qs] = T(0);
        X[qn] = T(0);
    }
}

// bcPhi with the periodic seam: k = km-1 is the copy of k = 0 (BoundaryConditions.h:422-438)
template<typename T>
__global__ void k_bc_phi(const BCLists<T> L, int im, int jm, int km, int pitch){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, f = blockIdx.z;
    if(j >= jm) return;
    const size_t r0 = ((size_t)i * jm + j) * pitch;
    T* X = L.phi[f];
    X[r0 + km - 1] = X[r0];
}

// bcSolidGround over the precomputed list of solid cells (BC_Jup.h:460-589). The extrapolation
// reads only fluid cells, which this pass does not write.
template<typename T>
__global__ void k_bc_solid(const BCLists<T> L, const SolidCell* __restrict__ cells, int n,
        int jm, int pitch, size_t plane, double c43, double c13){
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= n) return;
    const SolidCell c = cells[s];
    const size_t q = ((size_t)c.i * jm + c.j) * pitch + c.k;
    L.u[q] = T(0); L.v[q] = T(0); L.w[q] = T(0);
    const long R = pitch, P = (long)plane;
    long da = 0;
    switch(c.code){
        case 1: da = P; break;  case 2: da = -P; break;
        case 3: da = R; break;  case 4: da = -R; break;
        case 5: da = 1; break;  case 6: da = -1; break;
        default: da = 0; break;
    }
    if(da == 0){
        for(int f = 2; f < L.n_solid; f++) L.solid[f][q] = T(0);    // t, p_stat keep their values
        return;
    }
    for(int f = 0; f < L.n_solid; f++){
        T* X = L.solid[f];
        X[q] = (T)c43 * X[q + da] - (T)c13 * X[q + 2 * da];
    }
}

// MOUNT_T_AMBIENT (BC_Jup.h:593-608): the solid cells of a row take the mean t of its fluid cells.
// One block per row: the row is staged in LDS by all lanes, one lane sums it in k order (the
// reference's order, so the mean is bit-identical), all lanes write the solid cells.
constexpr int AMB_MAXK = 1024;
template<typename T>
__global__ __launch_bounds__(64) void k_bc_t_ambient(T* __restrict__ t, const uint8_t* __restrict__ msk,
        const int2* __restrict__ rows, int n, int jm, int km, int pitch){
    const int r = blockIdx.x;
    if(r >= n) return;
    using A = acc_t<T>;
    __shared__ A sv[AMB_MAXK];
    __shared__ uint8_t sm[AMB_MAXK];
    __shared__ A s_amb;
    __shared__ int s_cnt;
    const int2 ij = rows[r];
    const size_t r0 = ((size_t)ij.x * jm + ij.y) * pitch;
    for(int k = threadIdx.x; k < km; k += 64){ sv[k] = (A)t[r0 + k]; sm[k] = msk[r0 + k]; }
    __syncthreads();
    if(threadIdx.x == 0){
        // the sum in the reference's order; the LDS reads issued eight at a time ahead of the adds
        A s = 0.0; int cnt = 0;
        int k = 0;
        for(; k + 8 <= km; k += 8){
            A v[8]; uint8_t m[8];
            #pragma unroll
            for(int u = 0; u < 8; u++){ v[u] = sv[k + u]; m[u] = sm[k + u]; }
            #pragma unroll
            for(int u = 0; u < 8; u++) if(!m[u]){ s += v[u]; cnt++; }
        }
        for(; k < km; k++){
            if(sm[k]) continue;
            s += sv[k]; cnt++;
        }
        s_amb = (cnt > 0) ? s / (double)cnt : A(0.0);
        s_cnt = cnt;
    }
    __syncthreads();
    if(s_cnt == 0) return;
    const T t_amb = (T)s_amb;
    for(int k = threadIdx.x; k < km; k += 64) if(sm[k]) t[r0 + k] = t_amb;
}

// clampNegativeSpecies (FileIO_Jup.cpp:494-548)
template<typename T>
struct SpeciesPtrs {
    T* x[ATJUP_F_NH4SH - ATJUP_F_H2O + 1];
    float* lo[ATJUP_F_NH4SH - ATJUP_F_H2O + 1];     // the refined form's low words (else null)
};

template<typename T>
__global__ __launch_bounds__(256) void k_clamp_species(const SpeciesPtrs<T> S, size_t n){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    T* __restrict__ x = S.x[blockIdx.y];
    float* __restrict__ lo = S.lo[blockIdx.y];
    if(lo){
        if(!((double)x[q] + (double)lo[q] >= 0.0)){ x[q] = T(0); lo[q] = 0.0f; }
        return;
    }
    const T v = x[q];
    if(!(v >= T(0))) x[q] = T(0);
}

// the fluid run of every column from the bottom (the levels computeHydrostaticPressure integrates,
// RungeKutta_Jup_Turb.cpp:119-259): i0 the first fluid level (im when the column is solid), i1
// the last level of the run that starts there
__global__ void k_col_range(const uint8_t* __restrict__ flags, int2* __restrict__ colr, int im, int jm, int km,
        int pitch, size_t plane){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    int i0 = 0;
    while(i0 < im && (flags[q0 + i0 * plane] & FL_SELF)) i0++;
    int i1 = i0;
    if(i0 < im)
        while(i1 + 1 < im && !(flags[q0 + (i1 + 1) * plane] & FL_SELF)) i1++;
    colr[(size_t)j * km + k] = make_int2(i0, i1);
}

// ---------------------------------------------------------------------------------------------
// Obstacle set-up: face flags, i_topography and wall viscosity (computeWallViscosity,
// InitValues_Jup.cpp:84-134), solid-cell classification for bcSolidGround.
// ---------------------------------------------------------------------------------------------
__global__ void k_mask_flags(const uint8_t* __restrict__ msk, uint8_t* __restrict__ flags,
        int im, int jm, int km, int pitch, size_t plane){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const int kw = (k == 0) ? km - 2 : ((k == km - 1) ? km - 2 : k - 1);
    const int ke = (k >= km - 2) ? 0 : k + 1;
    uint8_t f = 0;
    if(msk[q]) f |= FL_SELF;
    if(i + 1 < im && msk[q + plane]) f |= FL_IP;
    if(i > 0 && msk[q - plane]) f |= FL_IM;
    if(j + 1 < jm && msk[q + pitch]) f |= FL_JP;
    if(j > 0 && msk[q - pitch]) f |= FL_JM;
    if(msk[q - k + ke]) f |= FL_KP;
    if(msk[q - k + kw]) f |= FL_KM;
    flags[q] = f;
}

__global__ void k_itopo(const uint8_t* __restrict__ msk, int* __restrict__ itopo, int im, int jm, int km,
        int pitch, size_t plane){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    int s = 0;
    while(s < im && msk[q0 + s * plane]) ++s;
    itopo[(size_t)j * km + k] = s;
}

template<typename T>
__global__ void k_wall_nue(const uint8_t* __restrict__ msk, T* __restrict__ wn, int im, int jm, int km,
        int pitch, size_t plane, int layers, double val_per_ramp, double re_factor){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    wn[q] = T(0);
    if(msk[q]) return;
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
                if(!msk[((size_t)ii * jm + jj) * pitch + kk]) continue;
                const int ad = max(abs(di), max(abs(dj), abs(dk)));
                if(ad < d) d = ad;
                if(d <= 1) break;
            }
        }
    }
    if(d > layers) return;
    const double ramp = (double)(layers - d + 1) / (double)layers;
    wn[q] = (T)(re_factor * ramp);
}

__global__ void k_solid_codes(const uint8_t* __restrict__ msk, int8_t* __restrict__ code, int im, int jm,
        int km, int pitch, size_t plane){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    if(!msk[q]){ code[q] = -1; return; }
    auto inb = [&](int ii, int jj, int kk){ return ii >= 0 && ii < im && jj >= 0 && jj < jm && kk >= 0 && kk < km; };
    auto M = [&](int ii, int jj, int kk){ return msk[((size_t)ii * jm + jj) * pitch + kk] != 0; };
    auto solid = [&](int ii, int jj, int kk){ return !inb(ii, jj, kk) || M(ii, jj, kk); };
    auto fluid = [&](int ii, int jj, int kk){ return inb(ii, jj, kk) && !M(ii, jj, kk); };
    const bool surf = !solid(i + 1, j, k) || !solid(i - 1, j, k) || !solid(i, j + 1, k)
                   || !solid(i, j - 1, k) || !solid(i, j, k + 1) || !solid(i, j, k - 1);
    int8_t c = 0;
    if(surf){
        if(fluid(i + 1, j, k) && fluid(i + 2, j, k)) c = 1;
        else if(fluid(i - 1, j, k) && fluid(i - 2, j, k)) c = 2;
        else if(fluid(i, j + 1, k) && fluid(i, j + 2, k)) c = 3;
        else if(fluid(i, j - 1, k) && fluid(i, j - 2, k)) c = 4;
        else if(fluid(i, j, k + 1) && fluid(i, j, k + 2)) c = 5;
        else if(fluid(i, j, k - 1) && fluid(i, j, k - 2)) c = 6;
        else c = 7;
    }
    code[q] = c;
}

template<typename T>
__global__ __launch_bounds__(256) void k_nonfinite(const T* __restrict__ x, int* __restrict__ flag, size_t n){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    if(!isfinite(x[q])) atomicOr(flag, 1);
}

// =============================================================================================
// host launchers
// =============================================================================================
namespace {

template<typename T> T* F(atjup_ctx* c, int f){ return (T*)c->cur[f].p; }

template<typename T>
int do_brl(atjup_ctx* c){
    hipLaunchKernelGGL(k_brl_partial<T>, dim3(BRL_NCH, c->im), dim3(BRL_BLOCK), 0, c->stream,
        F<T>(c, ATJUP_F_T), F<T>(c, ATJUP_F_P_STAT), (const uint8_t*)c->flags.p,
        (const double*)c->tabs.p + c->im, (acc_t<T>*)c->brl_part.p,
        c->jm, c->km, c->pitch, c->c.g, c->c.r_mix * c->c.R_mix, c->c.t_ref);
    hipLaunchKernelGGL(k_brl_final<acc_t<T>>, dim3((c->im + 63) / 64), dim3(64), 0, c->stream,
        (const acc_t<T>*)c->brl_part.p, (acc_t<T>*)c->brl.p, c->im);
    return check_launch();
}

template<typename T>
int do_hydro(atjup_ctx* c){
    hipLaunchKernelGGL(k_hydrostatic<T>, dim3(nblk(c->km, 128), c->jm), dim3(128), 0, c->stream,
        F<T>(c, ATJUP_F_V), F<T>(c, ATJUP_F_W), (const int2*)c->colr.p,
        (const double*)c->tabs.p, (const double*)c->tabs.p + c->im + c->jm,
        F<T>(c, ATJUP_F_P_HYDRO), c->im, c->jm, c->km, c->pitch, c->plane, c->g.dr);
    return check_launch();
}

template<typename T>
int do_rk4_dry(atjup_ctx* c, const RkBuffers* rb = nullptr){
    StageArgs<T> a{};
    const double dt = c->c.dt;
    a.pd = F<T>(c, ATJUP_F_P_DYN); a.ph = F<T>(c, ATJUP_F_P_HYDRO); a.ps = F<T>(c, ATJUP_F_P_STAT);
    a.rho = F<T>(c, ATJUP_F_RHO_MIX); a.wn = F<T>(c, ATJUP_F_WALL_NUE);
    a.aux_u = F<T>(c, ATJUP_F_AUX_U); a.aux_v = F<T>(c, ATJUP_F_AUX_V); a.aux_w = F<T>(c, ATJUP_F_AUX_W);
    a.flags = (const uint8_t*)c->flags.p;
    a.geo = (const GeoIJ<geo_t<T>>*)c->geo.p;
    a.brl = (const acc_t<T>*)c->brl.p;
    a.im = c->im; a.jm = c->jm; a.km = c->km; a.pitch = c->pitch; a.plane = c->plane;
    a.nkb = (c->km + ST_BK - 1) / ST_BK; a.njb = (c->jm + ST_BJ - 1) / ST_BJ;
    a.nb = a.nkb * a.njb * c->im;
    a.dt = (T)dt; a.one_sixth = (T)(1.0 / 6.0);
    a.dt_d = dt;
    for(int f = 0; f < 4; f++) a.lo[f] = (float*)c->lo[f].p;
    a.inv_2dr = (T)(1.0 / (2.0 * c->g.dr)); a.inv_2dthe = (T)(1.0 / (2.0 * c->g.dthe));
    a.inv_2dphi = (T)(1.0 / (2.0 * c->g.dphi)); a.inv_dr2 = (T)(1.0 / (c->g.dr * c->g.dr));
    a.inv_dthe2 = (T)(1.0 / (c->g.dthe * c->g.dthe)); a.inv_dphi2 = (T)(1.0 / (c->g.dphi * c->g.dphi));
    const atjup_consts& C = c->c;
    a.cep = (T)(C.u_0 * C.u_0 / (C.cp_mix * C.t_ref));
    a.to_nd = (T)(1.0 / (C.r_mix * C.u_0 * C.u_0 * 1.0e-5));
    a.r_mix = (T)C.r_mix;
    a.c_m2o = (T)(1.0 * -2.0 * C.omega);
    a.c_p2o = (T)(1.0 * +2.0 * C.omega);
    a.Coriolis = (T)C.Coriolis; a.centrifugal = (T)C.centrifugal;
    a.cbuoy = -1.0 * 1.0 * 1.0 * C.buoyancy;
    a.g = C.g; a.rR = C.r_mix * C.R_mix; a.t_ref_d = C.t_ref; a.t_ref = (T)C.t_ref;
    a.kt = (T)(c->kn.t_diff_scale / (C.re * C.pr) + 0.0 / 0.9);
    a.visc0 = (T)(1.0 / C.re + 0.0);
    a.chi = c->obst_on ? (const double*)c->chi.p : nullptr;
    a.inv_tau = c->obst_on ? 1.0 / c->obst.tau : 0.0;
    for(int s = 0; s < 4; s++){
        a.stage = s;
        a.c_in = (T)((s == 0 || s == 1) ? 0.5 * dt : (s == 2 ? dt : 0.0));
        a.wgt = (T)((s == 0 || s == 3) ? 1.0 : 2.0);
        a.copy_boundary = (s <= 1);
        for(int f = 0; f < 4; f++){
            T* N = F<T>(c, f);
            T* X = (T*)c->stg[0][f].p;
            T* Y = (T*)c->stg[1][f].p;
            const T* in = (s == 0) ? N : (s == 1 ? X : (s == 2 ? Y : X));
            T* out = (s == 0) ? X : (s == 1 ? Y : (s == 2 ? X : N));
            a.in[f] = in; a.out[f] = out; a.nl[f] = N; a.acc[f] = (T*)c->acc[f].p;
        }
        if(rb){
            // the recording run of the adjoint: every stage input kept (Y1, Y2, Y3), the step's
            // result in its own buffers, the n level, its low words and aux_* of the context left
            // as they were (the RK buffers hold T values; RkBuffers carries them untyped)
            a.copy_boundary = 1;
            for(int f = 0; f < 4; f++) a.lo[f] = nullptr;
            for(int f = 0; f < 4; f++){
                a.in[f] = (s == 0) ? F<T>(c, f) : (const T*)rb->y[s - 1][f];
                a.out[f] = (s < 3) ? (T*)rb->y[s][f] : (T*)rb->nn[f];
            }
            a.aux_u = (T*)rb->aux[0]; a.aux_v = (T*)rb->aux[1]; a.aux_w = (T*)rb->aux[2];
        }
        hipEvent_t t0 = timer_begin(c);
        hipLaunchKernelGGL(k_stage_dry<T>, dim3(a.nb), dim3(ST_BK, ST_BJ), 0, c->stream, a, a.geo);
        timer_end(c, ATJUP_T_RK_STAGES, t0, 1);
        if(int rc = check_launch()) return rc;
    }
    return 0;
}

// The colour kernels of n sweeps, replayed from a HIP graph captured on first use (one per sweep
// count, relaxation factor and kind), so that the 2n dependent launches of a solve go to the
// device as one submission.
template<typename T, typename S>
int run_sweeps(atjup_ctx* c, int kind, int n, int ghost_first, double omega, T* x0, T* x1, const S* s0, const S* s1){
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const int hp = c->rb_hp;
    const size_t hn = (size_t)c->im * c->jm * hp;
    const uint8_t* f0 = (const uint8_t*)c->rb_flags.p;
    const uint8_t* f1 = f0 + hn;
    const PoisIJ<double>* pw = (const PoisIJ<double>*)c->pois.p;
    const int nm = (c->km - 1) / 2;
    const int njb = (c->jm - 2 + RB_BJ - 1) / RB_BJ;
    // Pairs of cells per thread where the ring is even: they halve the memory instructions per
    // cell. Row-aligned pairs (one row per wavefront, the row's weights through scalar loads) where
    // they fill at least 90% of the lanes, else pairs flattened over (row, pair). Measured on
    // MI355X, one colour pass: 41 x 181 x 361 (90 pairs per row) FP64 10.2 us flattened, 10.9
    // row-aligned, 11.2 one cell per thread; FP32 7.9 / 8.7 / 9.2; 81 x 361 x 721 (180 pairs)
    // FP64 75.7 row-aligned, 79.4 flattened, 87.5 one cell.
    const int wf_pair = (nm / 2 + RB_BM - 1) / RB_BM;
    const bool pair_type = std::is_same<T, double>::value || std::is_same<T, float>::value;
    const bool pairs_ok = (nm % 2 == 0) && pair_type;
    // the flattened passes find a thread's row through a float reciprocal (div_small): fewer than
    // 2^22 groups per level
    const bool flat_ok = (long long)(c->jm - 2) * (nm / 2) + 255 < (1LL << 22);
    const bool pairs_flat = pairs_ok && flat_ok && (double)(nm / 2) / (RB_BM * wf_pair) < 0.9;
    const bool pairs = pairs_ok && !pairs_flat;
    const int ntb = ((c->jm - 2) * (nm / 2) + 255) / 256;          // flattened pairs per level
    const int nb_flat = ntb * (c->im - 2);
    // FP32: four cells per thread (flattened) where the ring allows; one float4 per neighbour row.
    // 41 x 181 x 361: 7.4 us against 7.9 us for flattened pairs; FP64 quads (two 16-byte loads per
    // row) are slower than FP64 pairs (11.8 against 10.4 us).
    const bool quads = pairs_ok && flat_ok && nm % 4 == 0 && std::is_same<T, float>::value;
    const int ntb4 = ((c->jm - 2) * (nm / 4) + 255) / 256;
    const int nb_quad = ntb4 * (c->im - 2);
    const int nmb = pairs ? (nm / 2 + RB_BM - 1) / RB_BM : (nm + RB_BM - 1) / RB_BM;
    const int nib = pairs ? c->im - 2 : (c->im - 2 + RB_NI - 1) / RB_NI;
    const int nb_sw = nmb * njb * nib;
    auto launch_all = [&](){
        for(int s = 0; s < n; s++)
            for(int colour = 0; colour < 2; colour++){
                const int gh = (s > 0 || ghost_first) ? 1 : 0;
                if constexpr(std::is_same<T, double>::value || std::is_same<T, float>::value)
                    if(quads){
                        hipLaunchKernelGGL((k_rb_sweep2f<T, S, 4>), dim3(nb_quad), dim3(256), 0, c->stream,
                            x0, x1, s0, s1, f0, f1, pw, c->im, c->jm, hp, nm, ntb4, nb_quad, colour,
                            gh, omega, c43, c13, 1.0f / (float)(nm / 4));
                        continue;
                    } else if(pairs_flat){
                        hipLaunchKernelGGL((k_rb_sweep2f<T, S, 2>), dim3(nb_flat), dim3(256), 0, c->stream,
                            x0, x1, s0, s1, f0, f1, pw, c->im, c->jm, hp, nm, ntb, nb_flat, colour,
                            gh, omega, c43, c13, 1.0f / (float)(nm / 2));
                        continue;
                    } else if(pairs){
                        hipLaunchKernelGGL((k_rb_sweep2<T, S>), dim3(nb_sw), dim3(RB_BM, RB_BJ), 0, c->stream,
                            x0, x1, s0, s1, f0, f1, pw, c->im, c->jm, hp, nm, nmb, njb, nb_sw, colour,
                            gh, omega, c43, c13);
                        continue;
                    }
                hipLaunchKernelGGL((k_rb_sweep<T, S>), dim3(nb_sw), dim3(RB_BM, RB_BJ), 0, c->stream,
                    x0, x1, s0, s1, f0, f1, pw, c->im, c->jm, hp, nm, nmb, njb, nb_sw, colour,
                    gh, omega, c43, c13);
            }
    };
#if ATJUP_SWEEP_GRAPH
    // the graph's key: everything its launches take as arguments besides the colour and the sweep
    const void* ptrs[7] = { x0, x1, s0, s1, f0, f1, pw };
    const int dims[4] = { c->im, c->jm, hp, nm };
    hipGraphExec_t exec = nullptr;
    atjup_ctx::SweepGraph* slot = nullptr;
    for(auto& g : c->sweep_graphs)
        if(g.kind == kind && g.n == n && g.ghost_first == ghost_first && g.omega == omega){ slot = &g; break; }
    if(slot && std::equal(ptrs, ptrs + 7, slot->ptrs) && std::equal(dims, dims + 4, slot->dims))
        exec = (hipGraphExec_t)slot->exec;
    if(!exec){
        hipGraph_t graph = nullptr;
        ATJUP_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_all();
        const hipError_t le = hipGetLastError();
        ATJUP_HIP_CHECK(hipStreamEndCapture(c->stream, &graph));
        if(le != hipSuccess){ (void)hipGraphDestroy(graph); set_error(std::string("sweep capture: ") + hipGetErrorString(le)); return ATJUP_ERR_HIP; }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(ie != hipSuccess){ set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); return ATJUP_ERR_HIP; }
        atjup_ctx::SweepGraph g{ kind, n, ghost_first, omega, (void*)exec, {}, {} };
        std::copy(ptrs, ptrs + 7, g.ptrs);
        std::copy(dims, dims + 4, g.dims);
        if(slot){
            // the same sequence on other buffers: the old executable is no longer replayed (the
            // stream is ordered, so a replay still queued has been submitted before this point)
            ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
            (void)hipGraphExecDestroy((hipGraphExec_t)slot->exec);
            *slot = g;
        } else {
            c->sweep_graphs.push_back(g);
        }
    }
    ATJUP_HIP_CHECK(hipGraphLaunch(exec, c->stream));
    return 0;
#else
    launch_all();
    return check_launch();
#endif
}

template<typename T>
int do_pressure(atjup_ctx* c, int sweeps, double omega){
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    T* p = F<T>(c, ATJUP_F_P_DYN);
    hipLaunchKernelGGL(k_prep_projection<T>, dim3(nblk(c->km - 1, 128), c->jm - 2), dim3(128), 0, c->stream,
        F<T>(c, ATJUP_F_AUX_U), F<T>(c, ATJUP_F_AUX_V), F<T>(c, ATJUP_F_AUX_W),
        c->im, c->jm, c->km, c->pitch, c->plane, c43, c13);
    const int nkb = (c->km - 1 + 63) / 64;
    const int nb_src = nkb * (c->jm - 2) * (c->im - 2);
    const int hp = c->rb_hp;
    const size_t hn = (size_t)c->im * c->jm * hp;
    using I = acc_t<T>;                       // the iterate: FP64 in the FP64 and FP32 forms
    I* s0 = (I*)c->psrc.p;
    I* s1 = s0 + hn;
    I* x0 = (I*)c->rb_p.p;
    I* x1 = x0 + hn;
    const PoisIJ<double>* pw = (const PoisIJ<double>*)c->pois.p;
    hipLaunchKernelGGL((k_poisson_source<T, I>), dim3(nb_src), dim3(64), 0, c->stream,
        F<T>(c, ATJUP_F_AUX_U), F<T>(c, ATJUP_F_AUX_V), F<T>(c, ATJUP_F_AUX_W),
        (const uint8_t*)c->flags.p, (const uint8_t*)c->mask.p, pw,
        s0, s1, c->im, c->jm, c->km, c->pitch, c->plane, nb_src, hp,
        1.0 / (2.0 * c->g.dr), 1.0 / (2.0 * c->g.dthe), 1.0 / (2.0 * c->g.dphi),
        1.0 / c->g.dthe, 1.0 / c->g.dphi, c->kn.pdyn_cap);
    dim3 gfull(nblk(c->km, 128), c->jm, c->im), b128(128);
    hipLaunchKernelGGL((k_rb_pack<T, I>), gfull, b128, 0, c->stream, (const T*)p, x0, x1, c->jm, c->km, c->pitch, hp);
    if(int rc = check_launch()) return rc;
    const int nm = (c->km - 1) / 2;
    const bool refined = c->precision == ATJUP_PRECISION_FP32_REF;
    // the FP64 sweeps: all of them in the drop-in; the first one in the refined form, whose planes
    // are those of p_dyn at entry (PressureSolver.h:323-358)
    const int n64 = refined ? 1 : sweeps;
    hipEvent_t t0 = timer_begin(c);
    if(int rc = run_sweeps<I, I>(c, 0, n64, 0, omega, x0, x1, (const I*)s0, (const I*)s1)) return rc;
    timer_end(c, ATJUP_T_POISSON_SWEEPS, t0, 2 * n64);
    if constexpr(std::is_same<T, float>::value) if(refined && sweeps > 1){
        float* d0 = (float*)c->rb_d.p;
        float* d1 = d0 + hn;
        float* r0 = (float*)c->rb_r.p;
        float* r1 = r0 + hn;
        const uint8_t* f0 = (const uint8_t*)c->rb_flags.p;
        const uint8_t* f1 = f0 + hn;
        dim3 gm(nblk(nm, 128), c->jm - 2, 2 * (c->im - 2));
        hipLaunchKernelGGL(k_fill_float, dim3(nblk(2 * hn, 256)), dim3(256), 0, c->stream, d0, 2 * hn, 0.0f);
        int left = sweeps - 1;
        while(left > 0){
            const int n = std::min(left, c->refine_sweeps);
            left -= n;
            hipLaunchKernelGGL(k_rb_planes_i, dim3(nblk(nm, 128), c->jm, 2), b128, 0, c->stream, x0, x1,
                c->im, c->jm, hp, nm, c43, c13);
            hipLaunchKernelGGL(k_rb_planes_j, dim3(nblk(nm, 128), c->im, 2), b128, 0, c->stream, x0, x1,
                c->im, c->jm, hp, nm, c43, c13);
            hipLaunchKernelGGL(k_rb_residual, gm, b128, 0, c->stream, (const double*)x0, (const double*)x1,
                (const double*)s0, (const double*)s1, f0, f1, pw, r0, r1, c->im, c->jm, hp, nm);
            hipEvent_t t1 = timer_begin(c);
            if(int rc = run_sweeps<float, float>(c, 1, n, 1, omega, d0, d1, (const float*)r0, (const float*)r1)) return rc;
            timer_end(c, ATJUP_T_POISSON_SWEEPS, t1, 2 * n);
            hipLaunchKernelGGL(k_rb_add, gm, b128, 0, c->stream, x0, x1, d0, d1, c->im, c->jm, hp, nm);
            if(int rc = check_launch()) return rc;
        }
    }
    hipLaunchKernelGGL((k_rb_unpack<T, I>), gfull, b128, 0, c->stream, p, (const I*)x0, (const I*)x1,
        c->jm, c->km, c->pitch, hp);
    hipLaunchKernelGGL(k_poisson_planes_i<T>, dim3(nblk(c->km, 128), c->jm), dim3(128), 0, c->stream,
        p, c->im, c->jm, c->km, c->pitch, c->plane, c43, c13);
    hipLaunchKernelGGL(k_poisson_planes_j<T>, dim3(nblk(c->km, 128), c->im), dim3(128), 0, c->stream,
        p, c->im, c->jm, c->km, c->pitch, c->plane, c43, c13);
    hipLaunchKernelGGL(k_seam_copy<T>, dim3(nblk(c->jm, 64), c->im), dim3(64), 0, c->stream,
        p, c->im, c->jm, c->km, c->pitch);
    return check_launch();
}

template<typename T>
int do_shapiro(atjup_ctx* c, int field, int order, double strength, int passes, int axes){
    dim3 grid(nblk(c->km, 128), c->jm, c->im), blk(128);
    for(int ps = 0; ps < passes; ps++){
        if(order == 2 && axes == 7){                 // the three directions in one launch
            const int nkb = nblk(c->km, SH_BK), njb = nblk(c->jm, SH_BJ), nch = nblk(c->im, SH_L);
            hipLaunchKernelGGL(k_shapiro2_fused<T>, dim3(nkb * njb * nch), dim3(SH_BK, SH_BJ), 0, c->stream,
                (const T*)c->cur[field].p, (T*)c->scratch.p, (const int*)c->itopo.p, c->im, c->jm, c->km,
                c->pitch, c->plane, nkb, njb, nch, (T)(strength * 0.25));
            std::swap(c->cur[field].p, c->scratch.p);
            continue;
        }
        if(axes & 4){
            hipLaunchKernelGGL(k_shapiro_k<T>, grid, blk, 0, c->stream, (const T*)c->cur[field].p, (T*)c->scratch.p,
                (const int*)c->itopo.p, c->im, c->jm, c->km, c->pitch, order, strength);
            std::swap(c->cur[field].p, c->scratch.p);
        }
        if(axes & 2){
            hipLaunchKernelGGL(k_shapiro_j<T>, grid, blk, 0, c->stream, (const T*)c->cur[field].p, (T*)c->scratch.p,
                (const int*)c->itopo.p, c->im, c->jm, c->km, c->pitch, order, strength);
            std::swap(c->cur[field].p, c->scratch.p);
        }
        if(axes & 1){
            hipLaunchKernelGGL(k_shapiro_i<T>, grid, blk, 0, c->stream, (const T*)c->cur[field].p, (T*)c->scratch.p,
                (const int*)c->itopo.p, c->im, c->jm, c->km, c->pitch, c->plane, order, strength);
            std::swap(c->cur[field].p, c->scratch.p);
        }
    }
    // the refined form's pair: the filtered value is the new hi, so its low word starts over
    if(passes > 0 && axes && field < ATJUP_N_PROG && c->lo[field].p)
        hipLaunchKernelGGL(k_fill_float, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream,
            (float*)c->lo[field].p, c->n3, 0.0f);
    return check_launch();
}

template<typename T>
int do_mixture(atjup_ctx* c){
    hipLaunchKernelGGL(k_mixture_density<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream,
        F<T>(c, ATJUP_F_T), F<T>(c, ATJUP_F_P_STAT), F<T>(c, ATJUP_F_P_DYN), F<T>(c, ATJUP_F_RHO_MIX),
        c->n3, c->c.t_ref, c->kn.rho_pdyn, c->c.r_mix * c->c.u_0 * c->c.u_0 * 1.0e-5, c->c.R_mix);
    return check_launch();
}

template<typename T>
BCLists<T> bc_lists(atjup_ctx* c){
    BCLists<T> L{};
    const bool dry = c->kn.dry != 0;
    auto add = [&](T** arr, int& n, int f){
        if(arr == L.radius)
            L.radius_role[n] = f == ATJUP_F_U ? 1 : f == ATJUP_F_V ? 2 : f == ATJUP_F_W ? 3
                             : (f >= ATJUP_F_H2O && f <= ATJUP_F_NH4SH) ? 4 : 0;
        arr[n++] = (T*)live_ptr(c, f);
    };
    const int species[] = { ATJUP_F_H2O, ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, ATJUP_F_H2S, ATJUP_F_NH3,
        ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_CH4, ATJUP_F_CH4_CLOUD, ATJUP_F_CH4_ICE, ATJUP_F_NH4SH };
    const int turbf[] = { ATJUP_F_TKE, ATJUP_F_DIS, ATJUP_F_NUE, ATJUP_F_PROD, ATJUP_F_TKE_SOURCE, ATJUP_F_DIS_SOURCE };
    const int flux3[] = { ATJUP_F_MASSFLUX_H2S, ATJUP_F_MASSFLUX_NH3, ATJUP_F_MASSFLUX_NH4SH };
    for(int f : { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W }) add(L.radius, L.n_radius, f);
    for(int f : { ATJUP_F_T, ATJUP_F_U }) add(L.theta_ex, L.n_theta_ex, f);
    for(int f : { ATJUP_F_V, ATJUP_F_W }) add(L.theta_zero, L.n_theta_zero, f);
    for(int f : { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W }) add(L.phi, L.n_phi, f);
    for(int f : { ATJUP_F_T, ATJUP_F_P_STAT, ATJUP_F_P_DYN }) add(L.solid, L.n_solid, f);
    if(!dry){
        for(int f : species){ add(L.radius, L.n_radius, f); add(L.theta_ex, L.n_theta_ex, f);
                              add(L.phi, L.n_phi, f); add(L.species, L.n_species, f); add(L.solid, L.n_solid, f); }
        for(int f : flux3){ add(L.radius, L.n_radius, f); add(L.theta_zero, L.n_theta_zero, f);
                            add(L.phi, L.n_phi, f); add(L.solid, L.n_solid, f); }
        add(L.theta_zero, L.n_theta_zero, ATJUP_F_FLUXLIM_NH4SH);
        add(L.phi, L.n_phi, ATJUP_F_FLUXLIM_NH4SH);
        add(L.solid, L.n_solid, ATJUP_F_FLUXLIM_NH4SH);
        for(int f : turbf){ add(L.radius, L.n_radius, f); add(L.theta_ex, L.n_theta_ex, f); add(L.phi, L.n_phi, f); }
        // the chemistry's carried fields, which its kernels write on the interior only: the
        // diffusive fluxes, thermalmassflux and the reaction rates (held in chem_w, in the fields'
        // precision); the boundary values these passes leave reach the next call's smoothing
        for(int f : { ATJUP_F_DIFFLUX_H2S, ATJUP_F_DIFFLUX_NH3, ATJUP_F_DIFFLUX_NH4SH }){
            add(L.radius, L.n_radius, f); add(L.theta_zero, L.n_theta_zero, f);
            add(L.phi, L.n_phi, f); add(L.solid, L.n_solid, f);
        }
        add(L.radius, L.n_radius, ATJUP_F_THERMALMASSFLUX); add(L.theta_ex, L.n_theta_ex, ATJUP_F_THERMALMASSFLUX);
        add(L.phi, L.n_phi, ATJUP_F_THERMALMASSFLUX); add(L.solid, L.n_solid, ATJUP_F_THERMALMASSFLUX);
        if(c->chem_w.p)
            for(int r = 0; r < 3; r++){
                T* w = (T*)c->chem_w.p + (size_t)r * c->n3;
                L.radius_role[L.n_radius] = 0;
                L.radius[L.n_radius++] = w;
                L.theta_ex[L.n_theta_ex++] = w;
                L.phi[L.n_phi++] = w;
                L.solid[L.n_solid++] = w;
            }
    }
    L.t = F<T>(c, ATJUP_F_T); L.u = F<T>(c, ATJUP_F_U); L.v = F<T>(c, ATJUP_F_V); L.w = F<T>(c, ATJUP_F_W);
    return L;
}

template<typename T>
int do_bc(atjup_ctx* c){
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    const BCLists<T> L = bc_lists<T>(c);
    hipLaunchKernelGGL(k_bc_zonal_means<T>, dim3(c->jm, 4), dim3(64), 0, c->stream,
        L.v, L.w, (acc_t<T>*)c->zm.p, c->im, c->jm, c->km, c->pitch);
    hipLaunchKernelGGL(k_bc_radius<T>, dim3(nblk(c->km, 128), c->jm, L.n_radius), dim3(128), 0, c->stream,
        L, (const acc_t<T>*)c->zm.p, c->im, c->jm, c->km, c->pitch, c->plane, c43, c13,
        (float*)c->lo[ATJUP_F_V].p, (float*)c->lo[ATJUP_F_W].p);
    hipLaunchKernelGGL(k_bc_theta<T>, dim3(nblk(c->km, 128), c->im, L.n_theta_ex + L.n_theta_zero), dim3(128), 0,
        c->stream, L, c->im, c->jm, c->km, c->pitch);
    hipLaunchKernelGGL(k_bc_phi<T>, dim3(nblk(c->jm, 64), c->im, L.n_phi), dim3(64), 0, c->stream,
        L, c->im, c->jm, c->km, c->pitch);
    if(c->n_solids > 0)
        hipLaunchKernelGGL(k_bc_solid<T>, dim3(nblk(c->n_solids, 64)), dim3(64), 0, c->stream,
            L, (const SolidCell*)c->solids.p, c->n_solids, c->jm, c->pitch, c->plane, c43, c13);
    if(c->n_amb_rows > 0)
        hipLaunchKernelGGL(k_bc_t_ambient<T>, dim3(c->n_amb_rows), dim3(64), 0, c->stream,
            L.t, (const uint8_t*)c->mask.p, (const int2*)c->amb_rows.p, c->n_amb_rows, c->jm, c->km, c->pitch);
    if(int rc = check_launch()) return rc;
    if(!c->kn.dry) return launch_turb_sanitize(c);
    return 0;
}

template<typename T>
int do_clamp(atjup_ctx* c){
    if(c->kn.dry) return 0;
    SpeciesPtrs<T> S{};
    for(int f = ATJUP_F_H2O; f <= ATJUP_F_NH4SH; f++){
        S.x[f - ATJUP_F_H2O] = F<T>(c, f);
        S.lo[f - ATJUP_F_H2O] = (float*)c->lo[f].p;
    }
    hipLaunchKernelGGL(k_clamp_species<T>, dim3(nblk(c->n3, 256), ATJUP_F_NH4SH - ATJUP_F_H2O + 1), dim3(256), 0,
        c->stream, S, c->n3);
    return check_launch();
}

template<typename T>
int do_mask_setup(atjup_ctx* c){
    dim3 grid(nblk(c->km, 128), c->jm, c->im), blk(128);
    hipLaunchKernelGGL(k_mask_flags, grid, blk, 0, c->stream, (const uint8_t*)c->mask.p, (uint8_t*)c->flags.p,
        c->im, c->jm, c->km, c->pitch, c->plane);
    {
        const size_t hn = (size_t)c->im * c->jm * c->rb_hp;
        uint8_t* f0 = (uint8_t*)c->rb_flags.p;
        hipLaunchKernelGGL(k_rb_flags, grid, blk, 0, c->stream, (const uint8_t*)c->flags.p, f0, f0 + hn,
            c->jm, c->km, c->pitch, c->rb_hp);
    }
    hipLaunchKernelGGL(k_col_range, dim3(nblk(c->km, 128), c->jm), blk, 0, c->stream, (const uint8_t*)c->flags.p,
        (int2*)c->colr.p, c->im, c->jm, c->km, c->pitch, c->plane);
    hipLaunchKernelGGL(k_itopo, dim3(nblk(c->km, 128), c->jm), blk, 0, c->stream, (const uint8_t*)c->mask.p,
        (int*)c->itopo.p, c->im, c->jm, c->km, c->pitch, c->plane);
    const int layers = c->kn.wall_nue_layers;
    const double factor = c->kn.wall_nue_factor;
    if(factor > 0.0 && layers > 0)
        hipLaunchKernelGGL(k_wall_nue<T>, grid, blk, 0, c->stream, (const uint8_t*)c->mask.p, F<T>(c, ATJUP_F_WALL_NUE),
            c->im, c->jm, c->km, c->pitch, c->plane, layers, 0.0, factor / c->c.re);
    else
        ATJUP_HIP_CHECK(hipMemsetAsync(c->cur[ATJUP_F_WALL_NUE].p, 0, c->cur[ATJUP_F_WALL_NUE].bytes, c->stream));
    // solid-cell classification, compacted on the host once per mask
    int8_t* dcode = nullptr;
    ATJUP_HIP_CHECK(hipMalloc((void**)&dcode, c->n3));
    hipLaunchKernelGGL(k_solid_codes, grid, blk, 0, c->stream, (const uint8_t*)c->mask.p, dcode,
        c->im, c->jm, c->km, c->pitch, c->plane);
    std::vector<int8_t> code(c->n3);
    ATJUP_HIP_CHECK(hipMemcpyAsync(code.data(), dcode, c->n3, hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    ATJUP_HIP_CHECK(hipFree(dcode));
    std::vector<SolidCell> cells;
    std::vector<int2> rows;
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++){
            bool any = false;
            for(int k = 0; k < c->km; k++){
                const int8_t cd = code[c->idx(i, j, k)];
                if(cd < 0) continue;
                any = true;
                cells.push_back({ i, j, k, cd == 7 ? 0 : cd });
            }
            if(any) rows.push_back(make_int2(i, j));
        }
    // code 7 (surface cell without a usable pair) is treated like an interior cell; code 0 too
    c->n_solids = (int)cells.size();
    c->n_amb_rows = (int)rows.size();
    if(c->solids.p){ ATJUP_HIP_CHECK(hipFree(c->solids.p)); c->solids.p = nullptr; }
    if(c->amb_rows.p){ ATJUP_HIP_CHECK(hipFree(c->amb_rows.p)); c->amb_rows.p = nullptr; }
    if(!cells.empty()){
        ATJUP_HIP_CHECK(hipMalloc(&c->solids.p, cells.size() * sizeof(SolidCell)));
        ATJUP_HIP_CHECK(hipMemcpy(c->solids.p, cells.data(), cells.size() * sizeof(SolidCell), hipMemcpyHostToDevice));
        ATJUP_HIP_CHECK(hipMalloc(&c->amb_rows.p, rows.size() * sizeof(int2)));
        ATJUP_HIP_CHECK(hipMemcpy(c->amb_rows.p, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
    }
    return check_launch();
}

template<typename T>
int do_check_finite(atjup_ctx* c){
    int* flag = (int*)c->nonfinite.p;
    ATJUP_HIP_CHECK(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
    const int nf = c->kn.dry ? 4 : ATJUP_N_PROG;
    for(int f = 0; f < nf; f++)
        hipLaunchKernelGGL(k_nonfinite<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream, F<T>(c, f), flag, c->n3);
    int h = 0;
    ATJUP_HIP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    return h ? ATJUP_NONFINITE : 0;
}

}  // namespace

#define DISPATCH(fn, ...) ((c->precision == ATJUP_PRECISION_FP64) ? fn<double>(__VA_ARGS__) :          \
                          (c->precision == ATJUP_PRECISION_FP32_REF) ? fn<float>(__VA_ARGS__) : fn<Dual>(__VA_ARGS__))

int launch_buoy_ref_level(atjup_ctx* c){ return DISPATCH(do_brl, c); }
int launch_hydrostatic(atjup_ctx* c){ return DISPATCH(do_hydro, c); }
int launch_rk4(atjup_ctx* c){
    if(int rc = launch_buoy_ref_level(c)) return rc;
    if(int rc = launch_hydrostatic(c)) return rc;
    if(!c->kn.dry) return launch_rk4_full(c);
    return DISPATCH(do_rk4_dry, c);
}
int launch_pressure(atjup_ctx* c, int sweeps, double omega){ return DISPATCH(do_pressure, c, sweeps, omega); }
template<typename T>
int do_prep_projection(atjup_ctx* c){
    hipLaunchKernelGGL(k_prep_projection<T>, dim3(nblk(c->km - 1, 128), c->jm - 2), dim3(128), 0, c->stream,
        (T*)c->cur[ATJUP_F_AUX_U].p, (T*)c->cur[ATJUP_F_AUX_V].p, (T*)c->cur[ATJUP_F_AUX_W].p,
        c->im, c->jm, c->km, c->pitch, c->plane, 4.0 / 3.0, 1.0 / 3.0);
    return check_launch();
}
int launch_prep_projection(atjup_ctx* c){
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("prepareProjectionBoundaries alone: FP64 and FP32 forms"); return ATJUP_ERR_PROFILE; }
    return (c->precision == ATJUP_PRECISION_FP64) ? do_prep_projection<double>(c) : do_prep_projection<float>(c);
}
int launch_bc_after_step(atjup_ctx* c){ return DISPATCH(do_bc, c); }
int launch_shapiro(atjup_ctx* c, int field, int order, double strength, int passes, int axes){
    return DISPATCH(do_shapiro, c, field, order, strength, passes, axes);
}
int launch_mixture_density(atjup_ctx* c){ return DISPATCH(do_mixture, c); }
int launch_clamp_species(atjup_ctx* c){ return DISPATCH(do_clamp, c); }
int launch_mask_setup(atjup_ctx* c){ return DISPATCH(do_mask_setup, c); }
int launch_rk4_dry_record(atjup_ctx* c, const RkBuffers& rb){
    if(!c->kn.dry || (c->precision != ATJUP_PRECISION_FP64 && c->precision != ATJUP_PRECISION_FP32_REF)){
        set_error("the adjoint covers the dynamics profile in the FP64 and FP32-refined forms"); return ATJUP_ERR_PROFILE;
    }
    if(int rc = launch_buoy_ref_level(c)) return rc;
    if(int rc = launch_hydrostatic(c)) return rc;
    return (c->precision == ATJUP_PRECISION_FP64) ? do_rk4_dry<double>(c, &rb) : do_rk4_dry<float>(c, &rb);
}
int launch_check_finite(atjup_ctx* c){ return DISPATCH(do_check_finite, c); }

}  // namespace atjup
