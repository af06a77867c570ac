// Hand-written adjoint of the FP64 dynamics step (the dry profile): the transposes of the boundary
// passes, the Shapiro filter on p_dyn, the mixture density, the hydrostatic column integral, the
// per-level buoyancy mean, the RK4 step with its right-hand side, and the red-black pressure
// solve, each the exact transpose of the linearised forward operator at the forward state held in
// the context. Branches (the source clamp, the validity tests, the solid-face mirrors) are frozen
// at the forward's choice. Adjoint fields live beside the forward fields in the same layout; the
// scatter of a cell's stencil adjoint into its neighbours uses FP64 atomic adds.
//
// Field order and formulas follow the forward kernels of atjup_kernels.hip, which restate
// RHS_Jup_Turb.cpp:31-1457 (dry: t, u, v, w), RungeKutta_Jup_Turb.cpp:47-583, PressureSolver.h:
// 145-766, lib/Utils.cpp:221-453, BC_Jup.h:185-658 and BoundaryConditions.h:151-518.
#include "atjup_device.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

namespace atjup {

template<typename T>
__device__ __forceinline__ void aadd(T* p, double v){ if(v != 0.0) unsafeAtomicAdd(p, (T)v); }

// =============================================================================================
// boundary passes (bcRadius, bcTheta, bcPhi, bcSolidGround, MOUNT_T_AMBIENT), transposed
// =============================================================================================
// MOUNT_T_AMBIENT^T: the solid cells of a row carried the row's fluid mean
template<typename T>
__global__ __launch_bounds__(64) void ka_t_ambient(T* __restrict__ lt, const uint8_t* __restrict__ msk,
        const int2* __restrict__ rows, int n, int jm, int km, int pitch){
    const int r = blockIdx.x;
    if(r >= n) return;
    const int2 ij = rows[r];
    const size_t r0 = ((size_t)ij.x * jm + ij.y) * pitch;
    __shared__ double s_sum;
    __shared__ int s_cnt;
    if(threadIdx.x == 0){
        double s = 0.0; int cnt = 0;
        for(int k = 0; k < km; k++){
            if(msk[r0 + k]) s += lt[r0 + k];
            else cnt++;
        }
        s_sum = s; s_cnt = cnt;
    }
    __syncthreads();
    if(s_cnt == 0) return;
    const double share = s_sum / s_cnt;
    for(int k = threadIdx.x; k < km; k += 64){
        if(msk[r0 + k]) lt[r0 + k] = 0.0;
        else lt[r0 + k] += share;
    }
}

// bcSolidGround^T over the solid-cell list: velocities zeroed; t and p_dyn extrapolated from a
// fluid pair (code 1..6) or, for p_dyn, zeroed (code 0)
template<typename T>
__global__ void ka_bc_solid(T* __restrict__ lu, T* __restrict__ lv, T* __restrict__ lw,
        T* __restrict__ lt, T* __restrict__ lp, const SolidCell* __restrict__ cells, int n,
        int jm, int pitch, size_t plane, double c43, double c13){
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if(s >= n) return;
    const SolidCell c = cells[s];
    const size_t q = ((size_t)c.i * jm + c.j) * pitch + c.k;
    lu[q] = 0.0; lv[q] = 0.0; lw[q] = 0.0;
    const long R = pitch, P = (long)plane;
    long da = 0;
    switch(c.code){
        case 1: da = P; break;  case 2: da = -P; break;
        case 3: da = R; break;  case 4: da = -R; break;
        case 5: da = 1; break;  case 6: da = -1; break;
        default: da = 0; break;
    }
    if(da == 0){ lp[q] = 0.0; return; }
    T* X[2] = { lt, lp };
    for(int f = 0; f < 2; f++){
        const double l = X[f][q];
        X[f][q] = 0.0;
        aadd(&X[f][q + da], c43 * l);
        aadd(&X[f][q + 2 * da], -c13 * l);
    }
}

// bcPhi^T: the seam meridian was the copy of k = 0
template<typename T>
__global__ void ka_bc_phi(T* __restrict__ a0, T* __restrict__ a1, T* __restrict__ a2,
        T* __restrict__ a3, int jm, int km, int pitch){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(j >= jm) return;
    const size_t r0 = ((size_t)i * jm + j) * pitch;
    T* X[4] = { a0, a1, a2, a3 };
    for(int f = 0; f < 4; f++){
        X[f][r0] += X[f][r0 + km - 1];
        X[f][r0 + km - 1] = 0.0;
    }
}

// bcTheta^T: t, u were copied from j = 3 (jm-4) into rows 0..2 (jm-3..jm-1); v, w ramped 1/3, 2/3
// into rows 1, 2 and zeroed at the pole row
template<typename T>
__global__ void ka_bc_theta(T* __restrict__ lt, T* __restrict__ lu, T* __restrict__ lv,
        T* __restrict__ lw, int jm, int km, int pitch){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(k >= km) return;
    const size_t R = (size_t)pitch;
    const size_t qs = (size_t)i * jm * pitch + k;
    const size_t qn = qs + (size_t)(jm - 1) * R;
    T* ex[2] = { lt, lu };
    for(int f = 0; f < 2; f++){
        T* X = ex[f];
        X[qs + 3 * R] += X[qs] + X[qs + R] + X[qs + 2 * R];
        X[qs] = X[qs + R] = X[qs + 2 * R] = 0.0;
        X[qn - 3 * R] += X[qn] + X[qn - R] + X[qn - 2 * R];
        X[qn] = X[qn - R] = X[qn - 2 * R] = 0.0;
    }
    T* z[2] = { lv, lw };
    for(int f = 0; f < 2; f++){
        T* X = z[f];
        X[qs + 3 * R] += (1.0 / 3.0) * X[qs + R] + (2.0 / 3.0) * X[qs + 2 * R];
        X[qs] = X[qs + R] = X[qs + 2 * R] = 0.0;
        X[qn - 3 * R] += (1.0 / 3.0) * X[qn - R] + (2.0 / 3.0) * X[qn - 2 * R];
        X[qn] = X[qn - R] = X[qn - 2 * R] = 0.0;
    }
}

// bcRadius^T, first the zonal sums the top sponge's means gather their adjoint from:
// zs[0][j] = sum_k (lV[im-1] + 2/3 lV[im-2]), zs[1][j] = sum_k lV[im-3] / 3, the same for w at 2, 3
template<typename T>
__global__ void ka_bc_zonal_sums(const T* __restrict__ lv, const T* __restrict__ lw, double* __restrict__ zs,
        int im, int jm, int km, int pitch, size_t plane){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, sel = blockIdx.y;
    if(j >= jm) return;
    const T* X = (sel & 2) ? lw : lv;
    const size_t r1 = ((size_t)(im - 1) * jm + j) * pitch;
    double s = 0.0;
    if((sel & 1) == 0){
        for(int k = 0; k < km; k++) s += X[r1 + k] + (2.0 / 3.0) * X[r1 - plane + k];
    } else {
        for(int k = 0; k < km; k++) s += X[r1 - 2 * plane + k] * (1.0 / 3.0);
    }
    zs[(size_t)sel * jm + j] = s;
}

template<typename T>
__global__ void ka_bc_radius(T* __restrict__ lt, T* __restrict__ lu, T* __restrict__ lv,
        T* __restrict__ lw, const double* __restrict__ zs, int im, int jm, int km, int pitch, size_t plane,
        double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    const size_t qt = q0 + (size_t)(im - 1) * plane;
    // t: 2-point Neumann at both radial walls
    lt[q0 + plane] += c43 * lt[q0];
    lt[q0 + 2 * plane] += -c13 * lt[q0];
    lt[q0] = 0.0;
    lt[qt - plane] += c43 * lt[qt];
    lt[qt - 2 * plane] += -c13 * lt[qt];
    lt[qt] = 0.0;
    // u: zero at both walls
    lu[q0] = 0.0;
    lu[qt] = 0.0;
    // v, w: Neumann at the bottom, the top sponge with its zonal means
    const double inv_n = 1.0 / (km - 1);
    T* X[2] = { lv, lw };
    for(int f = 0; f < 2; f++){
        T* V = X[f];
        V[q0 + plane] += c43 * V[q0];
        V[q0 + 2 * plane] += -c13 * V[q0];
        V[q0] = 0.0;
        const double s2 = zs[(size_t)(2 * f) * jm + j], s3 = zs[(size_t)(2 * f + 1) * jm + j];
        const double l2 = V[qt - plane], l3 = V[qt - 2 * plane];
        V[qt - plane] = (1.0 / 3.0) * l2 + ((k < km - 1) ? s2 * inv_n : 0.0);
        V[qt - 2 * plane] = (2.0 / 3.0) * l3 + ((k < km - 1) ? s3 * inv_n : 0.0);
        V[qt] = 0.0;
    }
}

// =============================================================================================
// Shapiro order 2 on one field (k pass, then j, then i), transposed pass by pass
// =============================================================================================
template<typename T>
__global__ __launch_bounds__(256) void ka_shapiro_i(const T* __restrict__ lo, T* __restrict__ li,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, size_t plane, double co){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, n = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)n * jm + j) * pitch + k;
    const int i_surf = max(itopo[(size_t)j * km + k], 0);
    if(n < i_surf){ li[q] = lo[q]; return; }
    double self = 1.0 - 2.0 * co;
    if(n == i_surf) self += co;                  // the lower neighbour of the first fluid cell is itself
    if(n == im - 1) self += co;                  // the upper neighbour of the top cell is itself
    double acc = self * lo[q];
    if(n + 1 <= im - 1) acc += co * lo[q + plane];
    if(n - 1 >= i_surf) acc += co * lo[q - plane];
    li[q] = acc;
}

template<typename T>
__global__ __launch_bounds__(256) void ka_shapiro_j(const T* __restrict__ lo, T* __restrict__ li,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, double co){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t R = (size_t)pitch;
    const size_t q = ((size_t)i * jm + n) * pitch + k;
    auto fl = [&](int jj){ return i >= max(itopo[(size_t)jj * km + k], 0); };
    if(!fl(n)){ li[q] = lo[q]; return; }
    const int jm1 = max(n - 1, 0), jp1 = min(n + 1, jm - 1);
    double self = 1.0 - 2.0 * co;
    if(!fl(jm1) || jm1 == n) self += co;
    if(!fl(jp1) || jp1 == n) self += co;
    double acc = self * lo[q];
    if(n + 1 <= jm - 1 && fl(n + 1)) acc += co * lo[q + R];
    if(n - 1 >= 0 && fl(n - 1)) acc += co * lo[q - R];
    li[q] = acc;
}

template<typename T>
__global__ __launch_bounds__(256) void ka_shapiro_k(const T* __restrict__ lo, T* __restrict__ li,
        const int* __restrict__ itopo, int im, int jm, int km, int pitch, double co){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t r0 = ((size_t)i * jm + j) * pitch;
    if(k == km - 1){ li[r0 + k] = 0.0; return; }      // the seam input is not read
    const int* it = itopo + (size_t)j * km;
    auto fl = [&](int kx){ return i >= max(it[kx], 0); };
    // the output at the seam is the copy of the output at k = 0
    auto L = [&](int kx){ return lo[r0 + kx] + ((kx == 0) ? lo[r0 + km - 1] : 0.0); };
    const int km1 = (k == 0) ? km - 2 : k - 1, kp1 = (k == km - 2) ? 0 : k + 1;
    if(!fl(k)){ li[r0 + k] = L(k); return; }
    double self = 1.0 - 2.0 * co;
    if(!fl(km1)) self += co;
    if(!fl(kp1)) self += co;
    double acc = self * L(k);
    if(fl(kp1)) acc += co * L(kp1);
    if(fl(km1)) acc += co * L(km1);
    li[r0 + k] = acc;
}

// =============================================================================================
// computeMixtureDensity^T and the hydrostatic column integral^T
// =============================================================================================
template<typename T>
__global__ __launch_bounds__(256) void ka_mixture(const T* __restrict__ t, const T* __restrict__ ps,
        const T* __restrict__ pd, T* __restrict__ lrho, T* __restrict__ lt, T* __restrict__ lp,
        size_t n, double t_ref, double pdyn_w, double ptb, double R_mix){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    const double l = lrho[q];
    lrho[q] = 0.0;
    if(l == 0.0) return;
    const double T_ = t[q] * t_ref;
    if(!(T_ > 0.0)) return;
    const double p_pa = ((double)ps[q] + pdyn_w * (double)pd[q] * ptb) * 1.0e5;
    const double r = p_pa / (R_mix * T_);
    if(!isfinite(r)) return;
    lt[q] += l * (-r / (double)t[q]);
    lp[q] += l * (pdyn_w * ptb * 1.0e5 / (R_mix * T_));
}

template<typename T>
__global__ __launch_bounds__(256) void ka_hydrostatic(const T* __restrict__ v, const T* __restrict__ w,
        const int2* __restrict__ colr, const double* __restrict__ rad, const double* __restrict__ cor_j,
        T* __restrict__ lph, T* __restrict__ lv, T* __restrict__ lw,
        int im, int jm, int km, int pitch, size_t plane, double dr){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if(k >= km) return;
    const size_t q0 = (size_t)j * pitch + k;
    const int2 run = colr[(size_t)j * km + k];     // the column's fluid run from the bottom
    const int i0 = run.x, i1 = run.y;
    if(i0 >= im){ for(int i = 0; i < im; i++) lph[q0 + i * plane] = 0.0; return; }
    const double cor_ = cor_j[j];
    double G = 0.0, lf_next = 0.0;
    for(int i = i1; i >= i0; i--){
        // f(i) receives 0.5 dr G_i from step i and 0.5 dr G_{i+1} from step i+1
        double lf = lf_next;
        if(i > i0){
            G += lph[q0 + i * plane];
            lf += 0.5 * dr * G;
            lf_next = 0.5 * dr * G;
        } else {
            lf_next = 0.0;
        }
        const double vv = (double)v[q0 + i * plane], ww = (double)w[q0 + i * plane];
        lv[q0 + i * plane] += lf * (2.0 * vv / rad[i]);
        lw[q0 + i * plane] += lf * (2.0 * ww / rad[i] + cor_);
    }
    for(int i = 0; i < im; i++) lph[q0 + i * plane] = 0.0;
}

// =============================================================================================
// computeBuoyancyRefLevel^T: the level mean's adjoint (the sum of the cells' -cb * lambda_buoy,
// gathered by the RHS adjoint into gbrl) spread back onto t with the mean's weights
// =============================================================================================
// Two passes: ABRL_NCH chunks of (j, k) per level reduce to partial sums, then one thread per
// level adds its partials in chunk order.
constexpr int ABRL_BLOCK = 256;
constexpr int ABRL_NCH = 32;
template<typename T>
__global__ __launch_bounds__(ABRL_BLOCK) void ka_brl_reduce(const T* __restrict__ t, const T* __restrict__ ps,
        const uint8_t* __restrict__ flags, const double* __restrict__ sin_the, const T* __restrict__ gbrl,
        double* __restrict__ part, int jm, int km, int pitch, double g, double rR, double t_ref){
    const int i = blockIdx.y, ch = blockIdx.x;
    const int ncell = jm * km;
    const int per = (ncell + ABRL_NCH - 1) / ABRL_NCH;
    const int c0 = ch * per, c1 = min(ncell, c0 + per);
    double sl = 0.0, sw = 0.0;
    for(int c = c0 + threadIdx.x; c < c1; c += ABRL_BLOCK){
        const int j = c / km, k = c - j * km;
        const size_t q = ((size_t)i * jm + j) * pitch + k;
        sl += gbrl[q];
        if(flags[q] & FL_SELF) continue;
        const double tq = (double)t[q];
        if(!(tq > 0.0)) continue;
        const double b = g * (double)ps[q] / (rR * tq * t_ref);
        if(!isfinite(b)) continue;
        sw += sin_the[j];
    }
    __shared__ double sh[2][ABRL_BLOCK];
    sh[0][threadIdx.x] = sl; sh[1][threadIdx.x] = sw;
    __syncthreads();
    for(int o = ABRL_BLOCK / 2; o > 0; o >>= 1){
        if((int)threadIdx.x < o){ sh[0][threadIdx.x] += sh[0][threadIdx.x + o]; sh[1][threadIdx.x] += sh[1][threadIdx.x + o]; }
        __syncthreads();
    }
    if(threadIdx.x == 0){
        part[((size_t)i * ABRL_NCH + ch) * 2 + 0] = sh[0][0];
        part[((size_t)i * ABRL_NCH + ch) * 2 + 1] = sh[1][0];
    }
}

__global__ void ka_brl_final(const double* __restrict__ part, double* __restrict__ out, int im){
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= im) return;
    double sl = 0.0, sw = 0.0;
    for(int ch = 0; ch < ABRL_NCH; ch++){
        sl += part[((size_t)i * ABRL_NCH + ch) * 2 + 0];
        sw += part[((size_t)i * ABRL_NCH + ch) * 2 + 1];
    }
    out[2 * i] = sl;
    out[2 * i + 1] = sw;
}

template<typename T>
__global__ __launch_bounds__(256) void ka_brl_spread(const T* __restrict__ t, const T* __restrict__ ps,
        const uint8_t* __restrict__ flags, const double* __restrict__ sin_the, const double* __restrict__ red,
        T* __restrict__ gbrl, T* __restrict__ lt, int jm, int km, int pitch, double g, double rR, double t_ref){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    gbrl[q] = 0.0;
    const double lbrl = red[2 * i], W = red[2 * i + 1];
    if(lbrl == 0.0 || !(W > 0.0)) return;
    if(flags[q] & FL_SELF) return;
    const double tq = (double)t[q];
    if(!(tq > 0.0)) return;
    const double b = g * (double)ps[q] / (rR * tq * t_ref);
    if(!isfinite(b)) return;
    lt[q] += lbrl * sin_the[j] / W * (-b / tq);
}

// =============================================================================================
// The smooth drag region standing in for the obstacle in gradients with respect to its
// semi-axes: chi = 1 / (1 + exp(-(1 - s) / width)), s the ellipsoid's normalised squared distance
// in index space, and dchi/d(a_phi_deg, b_the_deg, c_r_km) for the gradient's reduction.
// =============================================================================================
constexpr int OBST_BLOCKS = 1024;

struct ObstGeom {
    double j0, k0;
    double a, b, c;              // semi-axes in cells (phi, theta) and levels (r)
    double da, db, dc;           // d(a, b, c) / d(a_phi_deg, b_the_deg, c_r_km)
    double width;
    int im, jm, km, pitch;
};

// chi at (i, j, k) and, through ds, the derivatives of s with respect to a, b, c (cells)
__device__ __forceinline__ double obst_chi(const ObstGeom& o, int i, int j, int k, double* ds){
    const int kk = (k == o.km - 1) ? 0 : k;                  // the seam is the copy of k = 0
    const double xk = (kk - o.k0) / o.a, xj = (j - o.j0) / o.b, xi = i / o.c;
    const double s = xk * xk + xj * xj + xi * xi;
    if(ds){
        ds[0] = -2.0 * xk * xk / o.a;
        ds[1] = -2.0 * xj * xj / o.b;
        ds[2] = -2.0 * xi * xi / o.c;
    }
    const double x = (1.0 - s) / o.width;
    return (x < -36.8) ? 0.0 : 1.0 / (1.0 + exp(-x));       // below ~1e-16 it is taken as 0
}

__global__ __launch_bounds__(256) void k_obstacle_chi(double* __restrict__ chi, const ObstGeom o){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= o.km) return;
    chi[((size_t)i * o.jm + j) * o.pitch + k] = obst_chi(o, i, j, k, nullptr);
}

// sum over cells of dL/dchi * dchi/dp, p = the three semi-axes in their own units; per-block
// partial sums, added on the host in block order
__global__ __launch_bounds__(256) void ka_obstacle_grad(const double* __restrict__ gchi, double* __restrict__ part,
        const ObstGeom o){
    double acc[3] = { 0.0, 0.0, 0.0 };
    const size_t n = (size_t)o.im * o.jm * o.km;
    for(size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x){
        const int k = (int)(c % o.km);
        const size_t r = c / o.km;
        const int j = (int)(r % o.jm), i = (int)(r / o.jm);
        const double g = gchi[((size_t)i * o.jm + j) * o.pitch + k];
        if(g == 0.0) continue;
        double ds[3];
        const double x = obst_chi(o, i, j, k, ds);
        if(x == 0.0) continue;
        const double dchi_ds = -x * (1.0 - x) / o.width;
        acc[0] += g * dchi_ds * ds[0] * o.da;
        acc[1] += g * dchi_ds * ds[1] * o.db;
        acc[2] += g * dchi_ds * ds[2] * o.dc;
    }
    __shared__ double sh[3][256];
    for(int p = 0; p < 3; p++) sh[p][threadIdx.x] = acc[p];
    __syncthreads();
    for(int s = 128; s > 0; s >>= 1){
        if((int)threadIdx.x < s)
            for(int p = 0; p < 3; p++) sh[p][threadIdx.x] += sh[p][threadIdx.x + s];
        __syncthreads();
    }
    if(threadIdx.x == 0)
        for(int p = 0; p < 3; p++) part[(size_t)blockIdx.x * 3 + p] = sh[p][0];
}

static ObstGeom obst_geom(const atjup_ctx* c){
    const double deg = M_PI / 180.0;
    ObstGeom o{};
    o.j0 = c->obst.j0; o.k0 = c->obst.k0;
    o.da = deg / c->g.dphi;
    o.db = deg / c->g.dthe;
    o.dc = 1.0 / (c->g.L_atm * c->g.dr);
    o.a = c->obst.a_phi_deg * o.da;
    o.b = c->obst.b_the_deg * o.db;
    o.c = c->obst.c_r_km * o.dc;
    o.width = c->obst.width;
    o.im = c->im; o.jm = c->jm; o.km = c->km; o.pitch = c->pitch;
    return o;
}

// =============================================================================================
// RHS of the dry stage, transposed (RHS_Jup_Turb.cpp:31-1457 as restated by k_stage_dry)
// =============================================================================================
template<typename T>
struct RhsAdjArgs {
    const T* y[4];               // the stage input t, u, v, w
    const T* lk[4];              // the adjoint of the stage's right-hand side (interior cells)
    const T* la[3];              // the adjoint of aux_u, aux_v, aux_w (stage 3) or nullptr
    T* g[4];                     // the adjoint of the stage input (scatter target)
    const T *pd, *ph, *ps, *rho, *wn;
    T *lpd, *lph, *lrho, *gbrl;
    const uint8_t* flags;
    const GeoIJ<geo_t<T>>* geo;
    int im, jm, km, pitch, nkb, njb, nb;
    size_t plane;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dr2, inv_dthe2, inv_dphi2;
    double cep, to_nd, r_mix, c_m2o, c_p2o, Coriolis, cbuoy, g_, rR, t_ref, kt, visc0;
    const double* chi;           // the smooth drag region's indicator (else null)
    double inv_tau;
    double* gchi;                // dL/dchi, accumulated over stages and steps
};

// The scatter targets of one cell's stencil adjoints: its own cell and its six neighbours. The
// workgroup's (j, k) tile with a one-cell rim accumulates in LDS whatever lands in it (the cell,
// its j and k neighbours); the neighbours at i +- 1 and across the periodic seam go straight to
// global memory. The tile is added to global memory once, after every cell of the workgroup.
constexpr int RD_BK = 64, RD_BJ = 4;
constexpr int RD_NF = 6;                     // the scatter targets: g[0..3], lpd, lph
struct RhsTile { double v[RD_NF][RD_BJ + 2][RD_BK + 2]; };

template<typename T>
__device__ __forceinline__ void rhs_put(RhsTile& s, T* X, int f, size_t qg, int ty, int tx, double v){
    if(v == 0.0) return;
    if(ty >= 0) atomicAdd(&s.v[f][ty][tx], v);
    else unsafeAtomicAdd(X + qg, (T)v);
}

template<typename T>
__device__ __forceinline__ void rhs_dry_cell(const RhsAdjArgs<T>& a, const GeoIJ<geo_t<T>>* __restrict__ geo,
        RhsTile& tile, int i, int j, int k, size_t q, uint8_t fl){
    const size_t P = a.plane, R = (size_t)a.pitch;
    const int kw = (k == 0) ? a.km - 2 : k - 1;
    const int ke = (k == a.km - 2) ? 0 : k + 1;
    const size_t qkw = q - k + kw, qke = q - k + ke;
    // tile positions (row, column) of the cell and its in-tile neighbours; -1: global
    const int ty = (int)threadIdx.y + 1, tx = (int)threadIdx.x + 1;
    const int tx_e = (ke == k + 1) ? tx + 1 : -1, tx_w = (kw == k - 1) ? tx - 1 : -1;
    const GeoIJ<geo_t<T>> G = geo[wave_uniform(i * a.jm + j)];    // a wavefront is one (i, j) row

    double c_[4], d1r[4], d1t[4], d1p[4];
    for(int f = 0; f < 4; f++){
        const T* X = a.y[f];
        c_[f] = X[q];
        d1r[f] = ((double)X[q + P] - (double)X[q - P]) * a.inv_2dr;
        d1t[f] = ((double)X[q + R] - (double)X[q - R]) * a.inv_2dthe;
        d1p[f] = ((double)X[qke] - (double)X[qkw]) * a.inv_2dphi;
    }
    const double pdc = a.pd[q];
    const double dpdr   = (((fl & FL_IP) ? pdc : (double)a.pd[q + P]) - ((fl & FL_IM) ? pdc : (double)a.pd[q - P])) * a.inv_2dr;
    const double dpdthe = (((fl & FL_JP) ? pdc : (double)a.pd[q + R]) - ((fl & FL_JM) ? pdc : (double)a.pd[q - R])) * a.inv_2dthe;
    const double dpdphi = (((fl & FL_KP) ? pdc : (double)a.pd[qke]) - ((fl & FL_KM) ? pdc : (double)a.pd[qkw])) * a.inv_2dphi;
    const double tq = c_[0], u = c_[1], v = c_[2], w = c_[3];
    const double costhe = G.costhe, st = G.sinthe_true;
    const double v_invrm = v * G.inv_rm, w_invrs = w * G.inv_rmsinthe;
    const double S0 = ((double)a.ps[q + P] - (double)a.ps[q - P]) * a.inv_2dr * a.to_nd;
    const double rho_l = a.rho[q];
    const bool rvalid = rho_l > 0.0 && isfinite(rho_l);
    const double dpstatdr = rvalid ? S0 * (a.r_mix / rho_l) : S0;
    const double visc = a.visc0 + (double)a.wn[q];
    const double two_inv_rm = 2.0 * G.inv_rm;

    // adjoints of the four right-hand sides (with aux at the last stage)
    const double Lt = a.lk[0][q];
    const double lau = a.la[0] ? (double)a.la[0][q] : 0.0, lav = a.la[1] ? (double)a.la[1][q] : 0.0;
    const double law = a.la[2] ? (double)a.la[2][q] : 0.0;
    const double Lu = (double)a.lk[1][q] + lau, Lv = (double)a.lk[2][q] + lav, Lw = (double)a.lk[3][q] + law;

    double L_dpdr = -Lu + lau;
    double L_dpdthe = (-Lv + lav) * G.inv_rm, L_dphdthe = -Lv * G.inv_rm;
    double L_dpdphi = (-Lw + law) * G.inv_rmsinthe, L_dphdphi = -Lw * G.inv_rmsinthe;
    double Lc[4] = { 0.0, 0.0, 0.0, 0.0 };                         // centre-value adjoints
    double L1r[4], L1t[4], L1p[4], L2r[4], L2t[4], L2p[4];
    double L_vinv = 0.0, L_winv = 0.0;
    if(a.chi){
        // the drag region -chi / tau (u, v, w): its transpose, and dL/dchi of this stage
        const double pen = a.chi[q] * a.inv_tau;
        Lc[1] += -pen * Lu;
        Lc[2] += -pen * Lv;
        Lc[3] += -pen * Lw;
        a.gchi[q] += -(u * Lu + v * Lv + w * Lw) * a.inv_tau;
    }
    // transport
    const double Ltr[4] = { -Lt, -Lu, -Lv, -Lw };
    for(int f = 0; f < 4; f++){
        Lc[1] += Ltr[f] * d1r[f];
        L1r[f] = Ltr[f] * u;
        L_vinv += Ltr[f] * d1t[f];
        L1t[f] = Ltr[f] * v_invrm;
        L_winv += Ltr[f] * d1p[f];
        L1p[f] = Ltr[f] * w_invrs;
    }
    Lc[1] += Ltr[2] * v * G.inv_rm;
    Lc[2] += Ltr[2] * u * G.inv_rm;
    Lc[3] += Ltr[2] * (-2.0 * w * G.cotanthe * G.inv_rm);
    Lc[3] += Ltr[3] * (u + v * G.cotanthe) * G.inv_rm;
    Lc[1] += Ltr[3] * w * G.inv_rm;
    Lc[2] += Ltr[3] * w * G.cotanthe * G.inv_rm;
    // pressure work in the energy equation
    {
        const double Pt = a.cep * Lt;
        Lc[1] += Pt * (dpdr + dpstatdr);
        L_dpdr += Pt * u;
        const double L_dpstatdr = Pt * u;
        L_vinv += Pt * dpdthe;
        L_dpdthe += Pt * v_invrm;
        L_winv += Pt * dpdphi;
        L_dpdphi += Pt * w_invrs;
        if(rvalid) a.lrho[q] += L_dpstatdr * S0 * a.r_mix * (-1.0 / (rho_l * rho_l));
    }
    // diffusion
    const double Ld[4] = { a.kt * Lt, visc * Lu, visc * Lv, visc * Lw };
    for(int f = 0; f < 4; f++){
        L2r[f] = Ld[f];
        L2t[f] = Ld[f] * G.inv_rm2;
        L2p[f] = Ld[f] * G.inv_rm2sinthe2;
        L1t[f] += Ld[f] * G.costhe_inv_rm2sinthe;
    }
    L1r[0] += Ld[0] * two_inv_rm;
    Lc[1] += Ld[1] * 2.0 * G.inv_rm2;
    L1r[1] += Ld[1] * 4.0 * G.inv_rm;
    L1r[2] += Ld[2] * two_inv_rm;
    Lc[2] += -Ld[2] * G.v_metric;
    L1t[1] += Ld[2] * 2.0 * G.inv_rm2;
    L1p[3] += -Ld[2] * 2.0 * costhe * G.inv_rm2sinthe2;
    L1r[3] += Ld[3] * two_inv_rm;
    Lc[3] += -Ld[3] * G.v_metric;
    L1p[1] += Ld[3] * 2.0 * G.inv_rm2sinthe;
    L1p[2] += Ld[3] * 2.0 * costhe * G.inv_rm2sinthe2;
    // Coriolis
    Lc[3] += (-a.Coriolis * Lv) * a.c_m2o * costhe;
    {
        const double Lcp = -a.Coriolis * Lw;
        Lc[2] += Lcp * a.c_p2o * costhe;
        Lc[1] += Lcp * a.c_p2o * st;
    }
    Lc[2] += L_vinv * G.inv_rm;
    Lc[3] += L_winv * G.inv_rmsinthe;
    // buoyancy anomaly
    if(tq > 0.0){
        const double Lb = Lu;
        Lc[0] += Lb * a.cbuoy * (-a.g_ * (double)a.ps[q] / (a.rR * tq * tq * a.t_ref));
        a.gbrl[q] += -a.cbuoy * Lb;
    }
    // the stencils: the i +- 1 targets straight to global memory, the rest through the tile
    for(int f = 0; f < 4; f++){
        T* X = a.g[f];
        rhs_put(tile, X, f, q + P, -1, 0, a.inv_2dr * L1r[f] + a.inv_dr2 * L2r[f]);
        rhs_put(tile, X, f, q - P, -1, 0, -a.inv_2dr * L1r[f] + a.inv_dr2 * L2r[f]);
        rhs_put(tile, X, f, q + R, ty + 1, tx, a.inv_2dthe * L1t[f] + a.inv_dthe2 * L2t[f]);
        rhs_put(tile, X, f, q - R, ty - 1, tx, -a.inv_2dthe * L1t[f] + a.inv_dthe2 * L2t[f]);
        rhs_put(tile, X, f, qke, tx_e >= 0 ? ty : -1, tx_e, a.inv_2dphi * L1p[f] + a.inv_dphi2 * L2p[f]);
        rhs_put(tile, X, f, qkw, tx_w >= 0 ? ty : -1, tx_w, -a.inv_2dphi * L1p[f] + a.inv_dphi2 * L2p[f]);
        rhs_put(tile, X, f, q, ty, tx, -2.0 * (a.inv_dr2 * L2r[f] + a.inv_dthe2 * L2t[f] + a.inv_dphi2 * L2p[f]) + Lc[f]);
    }
    // the mirrored pressure gradients (a face towards a solid cell reads the cell itself)
    auto mir = [&](T* L, int f, double coef, uint8_t hi_flag, uint8_t lo_flag, size_t qh, size_t ql,
                   int hy, int hx, int ly, int lx){
        if(fl & hi_flag) rhs_put(tile, L, f, q, ty, tx, coef);
        else rhs_put(tile, L, f, qh, hy, hx, coef);
        if(fl & lo_flag) rhs_put(tile, L, f, q, ty, tx, -coef);
        else rhs_put(tile, L, f, ql, ly, lx, -coef);
    };
    const int ey = tx_e >= 0 ? ty : -1, wy = tx_w >= 0 ? ty : -1;
    mir(a.lpd, 4, a.inv_2dr * L_dpdr, FL_IP, FL_IM, q + P, q - P, -1, 0, -1, 0);
    mir(a.lpd, 4, a.inv_2dthe * L_dpdthe, FL_JP, FL_JM, q + R, q - R, ty + 1, tx, ty - 1, tx);
    mir(a.lpd, 4, a.inv_2dphi * L_dpdphi, FL_KP, FL_KM, qke, qkw, ey, tx_e, wy, tx_w);
    mir(a.lph, 5, a.inv_2dthe * L_dphdthe, FL_JP, FL_JM, q + R, q - R, ty + 1, tx, ty - 1, tx);
    mir(a.lph, 5, a.inv_2dphi * L_dphdphi, FL_KP, FL_KM, qke, qkw, ey, tx_e, wy, tx_w);
}

template<typename T>
__global__ __launch_bounds__(RD_BK * RD_BJ) void ka_rhs_dry(const RhsAdjArgs<T> a, const GeoIJ<geo_t<T>>* __restrict__ geo){
    __shared__ RhsTile tile;
    const int b = xcd_remap(blockIdx.x, a.nb);
    const int kb = b % a.nkb;
    const int rest = b / a.nkb;
    const int jb = rest % a.njb;
    const int i = rest / a.njb;
    const int k = kb * RD_BK + threadIdx.x;
    const int j = jb * RD_BJ + threadIdx.y;
    const int tid = threadIdx.y * RD_BK + threadIdx.x;
    double* tv = &tile.v[0][0][0];
    constexpr int NT = RD_NF * (RD_BJ + 2) * (RD_BK + 2);
    for(int c = tid; c < NT; c += RD_BK * RD_BJ) tv[c] = 0.0;
    __syncthreads();
    if(k < a.km && j < a.jm){
        const size_t q = ((size_t)i * a.jm + j) * a.pitch + k;
        const uint8_t fl = a.flags[q];
        const bool interior = i >= 1 && i <= a.im - 2 && j >= 3 && j <= a.jm - 4 && k <= a.km - 2 && !(fl & FL_SELF);
        if(interior) rhs_dry_cell(a, geo, tile, i, j, k, q, fl);
    }
    __syncthreads();
    // the tile and its rim into global memory, one add per touched cell
    constexpr int NC = (RD_BJ + 2) * (RD_BK + 2);
    for(int c = tid; c < RD_NF * NC; c += RD_BK * RD_BJ){
        const double v = tv[c];
        if(v == 0.0) continue;
        const int f = c / NC, r = c - f * NC;
        const int jj = jb * RD_BJ + r / (RD_BK + 2) - 1, kk = kb * RD_BK + r % (RD_BK + 2) - 1;
        if(jj < 0 || jj >= a.jm || kk < 0 || kk >= a.km) continue;      // holds no contribution
        T* X = f < 4 ? a.g[f] : (f == 4 ? a.lpd : a.lph);
        unsafeAtomicAdd(X + ((size_t)i * a.jm + jj) * a.pitch + kk, (T)v);
    }
}

// the RK4 combination transposed: the adjoint of stage s's right-hand side from the step output's
// adjoint and the adjoint of the next stage's input; the stage inputs' adjoints gather into the
// n level's
template<typename T>
__global__ __launch_bounds__(256) void ka_rk_combine(T* __restrict__ lk0, T* __restrict__ lk1,
        T* __restrict__ lk2, T* __restrict__ lk3, T* __restrict__ g0, T* __restrict__ g1,
        T* __restrict__ g2, T* __restrict__ g3, const T* __restrict__ lnp0, const T* __restrict__ lnp1,
        const T* __restrict__ lnp2, const T* __restrict__ lnp3, T* __restrict__ ln0, T* __restrict__ ln1,
        T* __restrict__ ln2, T* __restrict__ ln3, const uint8_t* __restrict__ flags,
        int im, int jm, int km, int pitch, double wnp, double cg, int have_g, int write_lk){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const bool interior = i >= 1 && i <= im - 2 && j >= 3 && j <= jm - 4 && k <= km - 2 && !(flags[q] & FL_SELF);
    T* lk[4] = { lk0, lk1, lk2, lk3 };
    T* g[4] = { g0, g1, g2, g3 };
    const T* lnp[4] = { lnp0, lnp1, lnp2, lnp3 };
    T* ln[4] = { ln0, ln1, ln2, ln3 };
    for(int f = 0; f < 4; f++){
        const double gv = have_g ? (double)g[f][q] : 0.0;
        if(write_lk) lk[f][q] = interior ? (wnp * (double)lnp[f][q] + cg * gv) : 0.0;
        if(have_g){ ln[f][q] += gv; g[f][q] = 0.0; }
    }
}

// =============================================================================================
// Red-black pressure solve transposed (reference formulation: colour 0, colour 1, then the outer
// planes i, j and the seam after every sweep). Works on the split storage of the forward solve.
// =============================================================================================
// planes^T: seam, then j planes, then i planes; the plane adjoints move into rows 1, 2
__global__ void ka_rb_seam(double* __restrict__ l0, double* __restrict__ l1, int im, int jm, int hp, int nm){
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if(j >= jm) return;
    // the seam cell k = 2 nm has the colour of k = 0: same array, compact index nm -> 0
    const int c = (i + j) & 1;
    double* L = c ? l1 : l0;
    const size_t r = ((size_t)i * jm + j) * hp;
    L[r] += L[r + nm];
    L[r + nm] = 0.0;
}
__global__ void ka_rb_planes_j(double* __restrict__ l0, double* __restrict__ l1, int im, int jm, int hp, int nm,
        double c43, double c13){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y, side = blockIdx.z;
    if(m > nm) return;
    const size_t R = (size_t)hp;
    const int ja = side ? jm - 1 : 0, jb1 = side ? jm - 2 : 1, jb2 = side ? jm - 3 : 2;
    for(int c = 0; c < 2; c++){
        // c: colour array of the plane cell (i, ja, m); (i, jb1) is in the other array, (i, jb2) in the same
        double* same = c ? l1 : l0;
        double* other = c ? l0 : l1;
        const size_t base = (size_t)i * jm * hp + m;
        const double l = same[base + ja * R];
        same[base + ja * R] = 0.0;
        other[base + jb1 * R] += c43 * l;
        same[base + jb2 * R] += -c13 * l;
    }
}
__global__ void ka_rb_planes_i(double* __restrict__ l0, double* __restrict__ l1, int im, int jm, int hp, int nm,
        double c43, double c13){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, side = blockIdx.z;
    if(m > nm) return;
    const size_t P = (size_t)jm * hp;
    const int ia = side ? im - 1 : 0, ib1 = side ? im - 2 : 1, ib2 = side ? im - 3 : 2;
    for(int c = 0; c < 2; c++){
        double* same = c ? l1 : l0;
        double* other = c ? l0 : l1;
        const size_t base = (size_t)j * hp + m;
        const double l = same[base + ia * P];
        same[base + ia * P] = 0.0;
        other[base + ib1 * P] += c43 * l;
        same[base + ib2 * P] += -c13 * l;
    }
}

// The sweeps transposed in the form the forward kernel runs them (k_rb_sweep): from the second
// sweep on, a plane value a sweep cell x reads is formed from x's own old value and the old value
// of the other-colour cell y behind it (c43 x - c13 y, x at i = 1, y at i = 2 and likewise at
// the other three planes), and the stored planes are only rewritten after the last sweep (their
// transpose runs once, before the sweeps). A colour-c pass, cell x of colour c:
//     x' = (1 - omega) x + (omega / den_x) (sum_d w_d v_d - src_x),
// v_d the neighbour read in direction d (a stored value, or a formed plane value). With
// mu_x = omega lambda_x' / den_x (0 where the cell is skipped, den_x = 0):
//     lambda_x  = (1 - omega) lambda_x' + sum over x's formed planes of c43 w_d mu_x,
//     lambda_v += w_d mu_x for every stored value v that x read,
//     lambda_y += -c13 w_d mu_x for the cell y behind each formed plane,
//     lambda_src_x -= mu_x.
// In a sweep, colour 0 reads colour 1 before the colour-1 pass and colour 1 reads the updated
// colour 0, except through formed planes, which hold colour 0's values from before the sweep. So,
// with mu of colour 1 known, the colour-0 kernel adds the direct reads by colour 1 to lambda_0
// before applying its own pass, and the formed-plane terms of colour 1 after it; the colour-1
// kernel applies its own pass and adds the reads by colour 0 (direct and formed), then forms the
// next sweep's mu of colour 1. Each colour keeps its mu in its own array.
__device__ __forceinline__ bool rb_reads(uint8_t fl_s, int dir){
    // does sweep cell s (flags fl_s) read its neighbour in direction dir (0 +i, 1 -i, 2 +j, 3 -j, 4 +k, 5 -k)
    if(fl_s & FL_SELF) return true;
    const uint8_t bit[6] = { FL_IP, FL_IM, FL_JP, FL_JM, FL_KP, FL_KM };
    return !(fl_s & bit[dir]);
}

__device__ __forceinline__ double rb_den(uint8_t fl, const PoisIJ<double>& W){
    if(fl & FL_SELF) return W.denom;
    double den = 0.0;
    if(!(fl & FL_IP)) den += W.num1;
    if(!(fl & FL_IM)) den += W.num1;
    if(!(fl & FL_JP)) den += W.num2;
    if(!(fl & FL_JM)) den += W.num2;
    if(!(fl & FL_KP)) den += W.num3;
    if(!(fl & FL_KM)) den += W.num3;
    return den;
}

// Which of a cell's six neighbours read it in a sweep (a sweep cell, reading in the opposite
// direction): bit d for the neighbour in direction d (0 +i, 1 -i, 2 +j, 3 -j, 4 +k, 5 -k). Fixed
// by the mask, formed once per transposed solve for both colours.
constexpr int RBT_BM = 64, RBT_BJ = 4;
__global__ __launch_bounds__(RBT_BM * RBT_BJ) void ka_rbT_inmask(const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1,
        uint8_t* __restrict__ in0, uint8_t* __restrict__ in1, int im, int jm, int hp, int nm){
    const int m = blockIdx.x * RBT_BM + threadIdx.x, j = blockIdx.y * RBT_BJ + threadIdx.y;
    const int i = blockIdx.z >> 1, colour = blockIdx.z & 1;
    if(m >= nm || j >= jm) return;
    const uint8_t* fo = colour ? f0 : f1;        // the other colour's flags
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp, q = row + m;
    auto sweep = [&](int si, int sj){ return si >= 1 && si <= im - 2 && sj >= 1 && sj <= jm - 2; };
    uint8_t bits = 0;
    if(i + 1 < im && sweep(i + 1, j) && rb_reads(fo[q + P], 1)) bits |= 1;
    if(i >= 1 && sweep(i - 1, j) && rb_reads(fo[q - P], 0)) bits |= 2;
    if(j + 1 < jm && sweep(i, j + 1) && rb_reads(fo[q + R], 3)) bits |= 4;
    if(j >= 1 && sweep(i, j - 1) && rb_reads(fo[q - R], 2)) bits |= 8;
    if(sweep(i, j)){
        const int off = (colour + i + j) & 1, ring = 2 * nm, kk = 2 * m + off;
        if(rb_reads(fo[row + (((kk + 1) % ring) >> 1)], 5)) bits |= 16;
        if(rb_reads(fo[row + (((kk - 1 + ring) % ring) >> 1)], 4)) bits |= 32;
    }
    (colour ? in1 : in0)[q] = bits;
}

// mu of colour 1 for the last sweep: omega lambda / den on the sweep cells, 0 elsewhere
template<typename TS>
__global__ __launch_bounds__(RBT_BM * RBT_BJ) void ka_rbT_mu(const TS* __restrict__ lam, TS* __restrict__ mu,
        const uint8_t* __restrict__ fl_arr, const PoisIJ<double>* __restrict__ pw, int im, int jm, int hp, int nm,
        TS omega){
    const int m = blockIdx.x * RBT_BM + threadIdx.x, j = blockIdx.y * RBT_BJ + threadIdx.y, i = blockIdx.z;
    if(m >= nm || j >= jm) return;
    const size_t q = ((size_t)i * jm + j) * hp + m;
    TS v = 0.0;
    if(i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2){
        const TS den = rb_den(fl_arr[q], pw[(size_t)i * jm + j]);
        if(den > 0.0) v = omega * lam[q] / den;
    }
    mu[q] = v;
}

// The values one transposed cell update needs from outside its own arrays: the other colour's mu
// at the six neighbour positions and its flags at the four radial and meridional ones, and the
// coupling weights of the (i, j) row and of its four neighbour rows. A caller fills the members
// the cell's incoming-read bits, position and pass use; the rest stay zero.
template<typename TS>
struct RbtNb {
    TS mip, mim, mjp, mjm, mke, mkw;          // mu_o at i+1, i-1, j+1, j-1, k+1, k-1
    TS wip, wim, wjp, wjm;                    // num1 of rows i+1, i-1; num2 of rows j+1, j-1
    uint8_t fim, fip, fjm, fjp;                   // the other colour's flags at i-1, i+1, j-1, j+1
};

// one cell of one transposed colour pass (W: the own row's weights, used on sweep cells only)
template<typename TS>
__device__ __forceinline__ void rbT_cell(TS* __restrict__ lam, TS* __restrict__ mu,
        TS* __restrict__ lsrc, size_t q, uint8_t bits, uint8_t fl, const PoisIJ<double>& W, const RbtNb<TS>& n,
        int i, int j, int im, int jm, int colour, int ghost, TS omega, TS c43, TS c13){
    const bool interior = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    // stored-value reads of this cell by the other colour's sweep cells s: sum w mu_s
    TS dsum = 0.0;
    if(bits & 1) dsum += n.wip * n.mip;
    if(bits & 2) dsum += n.wim * n.mim;
    if(bits & 4) dsum += n.wjp * n.mjp;
    if(bits & 8) dsum += n.wjm * n.mjm;
    if(!interior){
        lam[q] += dsum;                              // a plane cell, read as stored in the first sweep
        return;
    }
    if(bits & 16) dsum += W.num3 * n.mke;
    if(bits & 32) dsum += W.num3 * n.mkw;
    // formed-plane reads by the other colour, for which this cell is the one behind the plane
    TS bsum = 0.0;
    if(ghost){
        if(i == 2 && rb_reads(n.fim, 1))      bsum += -c13 * n.wim * n.mim;
        if(i == im - 3 && rb_reads(n.fip, 0)) bsum += -c13 * n.wip * n.mip;
        if(j == 2 && rb_reads(n.fjm, 3))      bsum += -c13 * n.wjm * n.mjm;
        if(j == jm - 3 && rb_reads(n.fjp, 2)) bsum += -c13 * n.wjp * n.mjp;
    }
    // this cell's own pass
    const TS den = rb_den(fl, W);
    const bool skip = !(den > 0.0);
    TS t = lam[q];
    if(colour == 0) t += dsum;                     // colour 1 read colour 0 after its pass
    TS mu_x = 0.0, out = t;
    if(!skip){
        mu_x = (colour == 0) ? omega * t / den : mu[q];
        TS self = 0.0;
        if(ghost){
            if(i == 1 && rb_reads(fl, 1))      self += c43 * W.num1 * mu_x;
            if(i == im - 2 && rb_reads(fl, 0)) self += c43 * W.num1 * mu_x;
            if(j == 1 && rb_reads(fl, 3))      self += c43 * W.num2 * mu_x;
            if(j == jm - 2 && rb_reads(fl, 2)) self += c43 * W.num2 * mu_x;
        }
        out = (1.0 - omega) * t + self;
        lsrc[q] += -mu_x;
    }
    out += bsum;
    if(colour == 1) out += dsum;                   // colour 0 read colour 1 before colour 1's pass
    lam[q] = out;
    // colour 0: its mu for the colour-1 kernel; colour 1: its mu for the previous sweep
    mu[q] = (colour == 0) ? mu_x : (skip ? 0.0 : omega * out / den);
}

// Which neighbour rows a row of cells needs: the stored-value reads (bits, the union over the
// row's cells) and, in sweeps that form the outer planes, the cells behind a plane.
struct RbtNeed {
    bool ip, im_, jp, jm_;
};
__device__ __forceinline__ RbtNeed rbt_need(uint8_t bits, bool interior, int ghost, int i, int j, int im, int jm){
    const bool g = ghost && interior;
    return RbtNeed{(bits & 1) || (g && i == im - 3), (bits & 2) || (g && i == 2),
                   (bits & 4) || (g && j == jm - 3), (bits & 8) || (g && j == 2)};
}

// one colour pass of one sweep, transposed, over every cell of colour c (plane cells take the
// reads of stored planes, which only the first sweep of a solve makes: ghost = 0)
template<typename TS>
__global__ __launch_bounds__(RBT_BM * RBT_BJ) void ka_rbT_pass(TS* __restrict__ lam, TS* __restrict__ mu,
        const TS* __restrict__ mu_o, TS* __restrict__ lsrc, const uint8_t* __restrict__ fl_arr,
        const uint8_t* __restrict__ fo_arr, const uint8_t* __restrict__ inm, const PoisIJ<double>* __restrict__ pw,
        int im, int jm, int hp, int nm, int colour, int ghost, TS omega, TS c43, TS c13){
    const int m = blockIdx.x * RBT_BM + threadIdx.x, j = blockIdx.y * RBT_BJ + threadIdx.y, i = blockIdx.z;
    if(m >= nm || j >= jm) return;
    const bool interior = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    if(!interior && ghost) return;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp;
    const size_t q = row + m;
    const uint8_t bits = inm[q];
    const int ij = wave_uniform(i * jm + j);          // a wavefront is one (i, j) row
    const RbtNeed nd = rbt_need(bits, interior, ghost, i, j, im, jm);
    RbtNb<TS> n = {};
    if(nd.ip){ n.mip = mu_o[q + P]; n.wip = pw[ij + jm].num1; }
    if(nd.im_){ n.mim = mu_o[q - P]; n.wim = pw[ij - jm].num1; }
    if(nd.jp){ n.mjp = mu_o[q + R]; n.wjp = pw[ij + 1].num2; }
    if(nd.jm_){ n.mjm = mu_o[q - R]; n.wjm = pw[ij - 1].num2; }
    if(bits & 48){
        const int off = (colour + i + j) & 1, ring = 2 * nm, kk = 2 * m + off;
        if(bits & 16) n.mke = mu_o[row + (((kk + 1) % ring) >> 1)];
        if(bits & 32) n.mkw = mu_o[row + (((kk - 1 + ring) % ring) >> 1)];
    }
    if(ghost){
        if(i == 2) n.fim = fo_arr[q - P];
        if(i == im - 3) n.fip = fo_arr[q + P];
        if(j == 2) n.fjm = fo_arr[q - R];
        if(j == jm - 3) n.fjp = fo_arr[q + R];
    }
    PoisIJ<double> W = {};
    uint8_t fl = 0;
    if(interior){ W = pw[ij]; fl = fl_arr[q]; }
    rbT_cell(lam, mu, lsrc, q, bits, fl, W, n, i, j, im, jm, colour, ghost, omega, c43, c13);
}

// the same pass with two cells m, m+1 (m even) per thread and pair loads, flattened over
// (row, pair) in 256-thread blocks so that no lane idles at the end of a row; needs an even ring of
// compact cells
template<typename TS>
__global__ __launch_bounds__(256) void ka_rbT_pass2(TS* __restrict__ lam, TS* __restrict__ mu,
        const TS* __restrict__ mu_o, TS* __restrict__ lsrc, const uint8_t* __restrict__ fl_arr,
        const uint8_t* __restrict__ fo_arr, const uint8_t* __restrict__ inm, const PoisIJ<double>* __restrict__ pw,
        int im, int jm, int hp, int nm, int ntb, int colour, int ghost, TS omega, TS c43, TS c13){
    using V2 = typename Pair<TS>::V;
    // levels fastest, each XCD a contiguous range: a group's neighbours at i +- 1 are read by the
    // workgroups dispatched beside it
    const int b = xcd_remap(blockIdx.x, ntb * im);
    const int i = b % im, tb = b / im;
    const int npair = nm / 2;
    const int t = tb * 256 + threadIdx.x;
    const int j = t / npair;
    const int m = 2 * (t - j * npair);
    if(j >= jm) return;
    const bool interior = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    if(!interior && ghost) return;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp;
    const size_t q = row + m;
    const uint16_t bits2 = *reinterpret_cast<const uint16_t*>(inm + q);
    const uint8_t b0 = (uint8_t)(bits2 & 0xff), b1 = (uint8_t)(bits2 >> 8);
    const int ij = i * jm + j;
    if(bits2 == 0x3f3f && i >= 3 && i <= im - 4 && j >= 3 && j <= jm - 4){
        // both cells fluid with six fluid neighbours that all read them, away from the rows that
        // form or sit behind the outer planes: rbT_cell without its branches, the pair at once
        const uint16_t fl2 = *reinterpret_cast<const uint16_t*>(fl_arr + q);
        if(fl2 == 0){
            const PoisIJ<double> W = pw[ij];
            const TS wip = pw[ij + jm].num1, wim = pw[ij - jm].num1;
            const TS wjp = pw[ij + 1].num2, wjm = pw[ij - 1].num2;
            const V2 vip = *reinterpret_cast<const V2*>(mu_o + q + P);
            const V2 vim = *reinterpret_cast<const V2*>(mu_o + q - P);
            const V2 vjp = *reinterpret_cast<const V2*>(mu_o + q + R);
            const V2 vjm = *reinterpret_cast<const V2*>(mu_o + q - R);
            const V2 vc = *reinterpret_cast<const V2*>(mu_o + q);
            const int off = (colour + i + j) & 1;
            const TS edge = off ? mu_o[row + (m + 2 == nm ? 0 : m + 2)] : mu_o[row + (m == 0 ? nm - 1 : m - 1)];
            TS mke[2], mkw[2];
            if(off){ mke[0] = vc.y; mkw[0] = vc.x; mke[1] = edge; mkw[1] = vc.y; }
            else   { mke[0] = vc.x; mkw[0] = edge; mke[1] = vc.y; mkw[1] = vc.x; }
            const TS mip[2] = { vip.x, vip.y }, mim[2] = { vim.x, vim.y };
            const TS mjp[2] = { vjp.x, vjp.y }, mjm[2] = { vjm.x, vjm.y };
            const TS den = W.den_fluid;
            const V2 lv = *reinterpret_cast<const V2*>(lam + q);
            const V2 sv = *reinterpret_cast<const V2*>(lsrc + q);
            V2 mv = colour ? *reinterpret_cast<const V2*>(mu + q) : V2{TS(0), TS(0)};
            const TS lin[2] = { lv.x, lv.y }, sin_[2] = { sv.x, sv.y };
            const TS min_[2] = { mv.x, mv.y };
            TS lout[2], sout[2], mout[2];
            #pragma unroll
            for(int e = 0; e < 2; e++){
                TS dsum = 0.0;
                dsum += wip * mip[e];
                dsum += wim * mim[e];
                dsum += wjp * mjp[e];
                dsum += wjm * mjm[e];
                dsum += W.num3 * mke[e];
                dsum += W.num3 * mkw[e];
                TS t = lin[e];
                if(colour == 0) t += dsum;
                const TS mu_x = (colour == 0) ? omega * t / den : min_[e];
                TS out = (1.0 - omega) * t + 0.0;
                sout[e] = sin_[e] + -mu_x;
                if(colour == 1) out += dsum;
                lout[e] = out;
                mout[e] = (colour == 0) ? mu_x : omega * out / den;
            }
            *reinterpret_cast<V2*>(lam + q) = V2{lout[0], lout[1]};
            *reinterpret_cast<V2*>(lsrc + q) = V2{sout[0], sout[1]};
            *reinterpret_cast<V2*>(mu + q) = V2{mout[0], mout[1]};
            return;
        }
    }
    const RbtNeed nd = rbt_need(b0 | b1, interior, ghost, i, j, im, jm);
    RbtNb<TS> n0 = {}, n1 = {};
    if(nd.ip){
        const V2 v = *reinterpret_cast<const V2*>(mu_o + q + P);
        n0.mip = v.x; n1.mip = v.y; n0.wip = n1.wip = pw[ij + jm].num1;
    }
    if(nd.im_){
        const V2 v = *reinterpret_cast<const V2*>(mu_o + q - P);
        n0.mim = v.x; n1.mim = v.y; n0.wim = n1.wim = pw[ij - jm].num1;
    }
    if(nd.jp){
        const V2 v = *reinterpret_cast<const V2*>(mu_o + q + R);
        n0.mjp = v.x; n1.mjp = v.y; n0.wjp = n1.wjp = pw[ij + 1].num2;
    }
    if(nd.jm_){
        const V2 v = *reinterpret_cast<const V2*>(mu_o + q - R);
        n0.mjm = v.x; n1.mjm = v.y; n0.wjm = n1.wjm = pw[ij - 1].num2;
    }
    // phi neighbours: the reader at k+1 and at k-1 of each cell (other colour)
    if((b0 | b1) & 48){
        const int off = (colour + i + j) & 1;
        const V2 vc = *reinterpret_cast<const V2*>(mu_o + q);
        const TS edge = off ? mu_o[row + (m + 2 == nm ? 0 : m + 2)] : mu_o[row + (m == 0 ? nm - 1 : m - 1)];
        if(off){ n0.mke = vc.y; n0.mkw = vc.x; n1.mke = edge; n1.mkw = vc.y; }
        else   { n0.mke = vc.x; n0.mkw = edge; n1.mke = vc.y; n1.mkw = vc.x; }
    }
    if(ghost){
        auto split = [&](const uint8_t* a, uint8_t& x0, uint8_t& x1){
            const uint16_t v = *reinterpret_cast<const uint16_t*>(a);
            x0 = (uint8_t)(v & 0xff); x1 = (uint8_t)(v >> 8);
        };
        if(i == 2) split(fo_arr + q - P, n0.fim, n1.fim);
        if(i == im - 3) split(fo_arr + q + P, n0.fip, n1.fip);
        if(j == 2) split(fo_arr + q - R, n0.fjm, n1.fjm);
        if(j == jm - 3) split(fo_arr + q + R, n0.fjp, n1.fjp);
    }
    PoisIJ<double> W = {};
    uint16_t fl2 = 0;
    if(interior){ W = pw[ij]; fl2 = *reinterpret_cast<const uint16_t*>(fl_arr + q); }
    rbT_cell(lam, mu, lsrc, q, b0, (uint8_t)(fl2 & 0xff), W, n0, i, j, im, jm, colour, ghost, omega, c43, c13);
    rbT_cell(lam, mu, lsrc, q + 1, b1, (uint8_t)(fl2 >> 8), W, n1, i, j, im, jm, colour, ghost, omega, c43, c13);
}

// ---- the FP32-refined form's rounds, transposed ----------------------------------------------
// A round of the refined solve forms the outer planes of the FP64 iterate x from its interior,
// the FP64 Gauss-Seidel residual r = src + den x - sum(w x_nb) of every sweep cell, relaxes the
// correction d from zero with n FP32 sweeps on the source r, and adds d to x on the sweep cells.
// Transposed, in reverse: the correction's adjoint is the iterate's on the sweep cells (x += d),
// the FP32 sweeps run transposed on it and leave the adjoint of r in their source adjoint, r's
// transpose adds into the FP64 adjoints of x and of the source, and the planes' transposes move
// the plane cells' adjoint into rows 1 and 2.

// x += d, transposed: lambda_d = lambda_x on the sweep cells (0 elsewhere); the round's lambda_r
// starts at zero
__global__ __launch_bounds__(128) void ka_rb_addT(const double* __restrict__ lx0, const double* __restrict__ lx1,
        float* __restrict__ ld0, float* __restrict__ ld1, float* __restrict__ lr0, float* __restrict__ lr1,
        int im, int jm, int hp, int nm){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const int i = blockIdx.z >> 1, colour = blockIdx.z & 1;
    if(m >= nm) return;
    const size_t q = ((size_t)i * jm + j) * hp + m;
    const bool sweep = i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2;
    (colour ? ld1 : ld0)[q] = sweep ? (float)(colour ? lx1 : lx0)[q] : 0.0f;
    (colour ? lr1 : lr0)[q] = 0.0f;
}

// r = src + den x - sum(w x_nb) over the sweep cells, transposed and gathered per cell y of either
// colour (plane cells included): lambda_x(y) += den_y lambda_r(y) - sum over the other colour's
// sweep cells z that read y of w_z lambda_r(z) (w_z the reader row's weight of that direction, as
// the residual applies it; the incoming-read bits of y list the readers); lambda_src(y) += lambda_r(y)
__global__ __launch_bounds__(128) void ka_rb_resT(double* __restrict__ lx0, double* __restrict__ lx1,
        double* __restrict__ ls0, double* __restrict__ ls1, const float* __restrict__ lr0, const float* __restrict__ lr1,
        const uint8_t* __restrict__ f0, const uint8_t* __restrict__ f1, const uint8_t* __restrict__ in0,
        const uint8_t* __restrict__ in1, const PoisIJ<double>* __restrict__ pw, int im, int jm, int hp, int nm){
    const int m = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    const int i = blockIdx.z >> 1, colour = blockIdx.z & 1;
    if(m >= nm) return;
    const size_t P = (size_t)jm * hp, R = (size_t)hp;
    const size_t row = ((size_t)i * jm + j) * hp, q = row + m;
    const float* lro = colour ? lr0 : lr1;              // the readers are of the other colour
    const uint8_t bits = (colour ? in1 : in0)[q];
    double g = 0.0;
    if(bits & 1) g -= pw[(size_t)(i + 1) * jm + j].num1 * (double)lro[q + P];
    if(bits & 2) g -= pw[(size_t)(i - 1) * jm + j].num1 * (double)lro[q - P];
    if(bits & 4) g -= pw[(size_t)i * jm + j + 1].num2 * (double)lro[q + R];
    if(bits & 8) g -= pw[(size_t)i * jm + j - 1].num2 * (double)lro[q - R];
    if(bits & 48){
        const int off = (colour + i + j) & 1, ring = 2 * nm, kk = 2 * m + off;
        const double w3 = pw[(size_t)i * jm + j].num3;
        if(bits & 16) g -= w3 * (double)lro[row + (((kk + 1) % ring) >> 1)];
        if(bits & 32) g -= w3 * (double)lro[row + (((kk - 1 + ring) % ring) >> 1)];
    }
    if(i >= 1 && i <= im - 2 && j >= 1 && j <= jm - 2){
        const double lr = (double)(colour ? lr1 : lr0)[q];
        g += rb_den((colour ? f1 : f0)[q], pw[(size_t)i * jm + j]) * lr;
        (colour ? ls1 : ls0)[q] += lr;
    }
    (colour ? lx1 : lx0)[q] += g;
}

// prepareProjectionBoundaries + the clamped divergence source, transposed: the source adjoint
// (frozen to zero where the forward clamped) spreads onto aux_u, aux_v, aux_w
template<typename T>
struct SrcAdjArgs {
    const T *au, *av, *aw;                  // forward aux after prepareProjectionBoundaries
    const double *ls0, *ls1;                // source adjoints, split (the source is FP64 in both forms)
    T *lau, *lav, *law;
    const uint8_t *flags, *msk;
    const PoisIJ<double>* pw;
    int im, jm, km, pitch, hp;
    size_t plane;
    double inv_2dr, inv_2dthe, inv_2dphi, inv_dthe, inv_dphi, cap;
};

template<typename T>
__global__ __launch_bounds__(256) void ka_poisson_source(const SrcAdjArgs<T> a){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + 1, i = blockIdx.z + 1;
    if(k > a.km - 2 || j > a.jm - 2 || i > a.im - 2) return;
    const int im = a.im, jm = a.jm, km = a.km;
    const size_t q = ((size_t)i * jm + j) * a.pitch + k;
    const size_t P = a.plane, R = (size_t)a.pitch;
    const double ls = (((i + j + k) & 1) ? a.ls1 : a.ls0)[((size_t)i * jm + j) * a.hp + (k >> 1)];
    if(ls == 0.0) return;
    const uint8_t fl = a.flags[q];
    const bool sld = fl & FL_SELF;
    const int kw = (k == 0) ? km - 2 : k - 1, ke = (k == km - 2) ? 0 : k + 1;
    const bool skp = fl & FL_KP, skm = fl & FL_KM;
    const bool skp2 = (k < km - 2) ? (a.msk[q + 2] != 0) : false;
    const bool skm2 = (k > 2) ? (a.msk[q - 2] != 0) : false;
    // forward value, to know whether the clamp held it
    double du_dr = 0.0, dv_dthe = 0.0, dw_dphi = 0.0;
    int r_mode = 0, t_mode = 0, p_mode = 0;     // which stencil each derivative took
    if(i < im - 2 && sld && !(fl & FL_IP)){
        du_dr = (-3.0 * (double)a.au[q] + 4.0 * (double)a.au[q + P] - (double)a.au[q + 2 * P]) * a.inv_2dr; r_mode = 1;
    }
    if(j > 2 && j < jm - 2){
        const bool air_jp1 = !(fl & FL_JP), air_jm1 = !(fl & FL_JM);
        const bool sjp2 = a.msk[q + 2 * R] != 0, sjm2 = a.msk[q - 2 * R] != 0;
        if(sld && air_jp1 && !sjp2){ dv_dthe = (-3.0 * (double)a.av[q] + 4.0 * (double)a.av[q + R] - (double)a.av[q + 2 * R]) * a.inv_2dthe; t_mode = 1; }
        if(sld && air_jm1 && !sjm2){ dv_dthe = -(-3.0 * (double)a.av[q] + 4.0 * (double)a.av[q - R] - (double)a.av[q - 2 * R]) * a.inv_2dthe; t_mode = 2; }
        if((sld && air_jp1 && sjp2) || (j == jm - 2 && !sld && (fl & FL_JP))){ dv_dthe = ((double)a.av[q + R] - (double)a.av[q]) * a.inv_dthe; t_mode = 3; }
        if((sld && air_jm1 && sjm2) || (j == 1 && sld && air_jm1)){ dv_dthe = ((double)a.av[q - R] - (double)a.av[q]) * a.inv_dthe; t_mode = 4; }
    }
    if(k > 2 && k < km - 2){
        const bool air_kp1 = !skp, air_km1 = !skm;
        if(sld && air_kp1 && !skp2){ dw_dphi = (-3.0 * (double)a.aw[q] + 4.0 * (double)a.aw[q + 1] - (double)a.aw[q + 2]) * a.inv_2dphi; p_mode = 1; }
        if(sld && air_km1 && !skm2){ dw_dphi = -(-3.0 * (double)a.aw[q] + 4.0 * (double)a.aw[q - 1] - (double)a.aw[q - 2]) * a.inv_2dphi; p_mode = 2; }
        if((sld && air_kp1 && skp2) || (k == km - 2 && !sld && skp)){ dw_dphi = ((double)a.aw[q + 1] - (double)a.aw[q]) * a.inv_dphi; p_mode = 3; }
        if((sld && air_km1 && skm2) || (k == 1 && sld && air_km1)){ dw_dphi = ((double)a.aw[q - 1] - (double)a.aw[q]) * a.inv_dphi; p_mode = 4; }
    }
    if(!r_mode) du_dr = ((double)a.au[q + P] - (double)a.au[q - P]) * a.inv_2dr;
    if(!t_mode) dv_dthe = ((double)a.av[q + R] - (double)a.av[q - R]) * a.inv_2dthe;
    if(!p_mode) dw_dphi = ((double)a.aw[q - k + ke] - (double)a.aw[q - k + kw]) * a.inv_2dphi;
    const PoisIJ<double> W = a.pw[(size_t)i * jm + j];
    const double div_src = du_dr * 1.0 + dv_dthe * W.inv_rm + dw_dphi * W.inv_rmsinthe;
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
    const double src_max = den * a.cap;
    if(!isfinite(div_src) || div_src > src_max || div_src < -src_max) return;    // clamped: frozen
    const double lr = ls * 1.0, lt = ls * W.inv_rm, lp = ls * W.inv_rmsinthe;
    switch(r_mode){
        case 1: aadd(&a.lau[q], -3.0 * a.inv_2dr * lr); aadd(&a.lau[q + P], 4.0 * a.inv_2dr * lr);
                aadd(&a.lau[q + 2 * P], -a.inv_2dr * lr); break;
        default: aadd(&a.lau[q + P], a.inv_2dr * lr); aadd(&a.lau[q - P], -a.inv_2dr * lr); break;
    }
    switch(t_mode){
        case 1: aadd(&a.lav[q], -3.0 * a.inv_2dthe * lt); aadd(&a.lav[q + R], 4.0 * a.inv_2dthe * lt);
                aadd(&a.lav[q + 2 * R], -a.inv_2dthe * lt); break;
        case 2: aadd(&a.lav[q], 3.0 * a.inv_2dthe * lt); aadd(&a.lav[q - R], -4.0 * a.inv_2dthe * lt);
                aadd(&a.lav[q - 2 * R], a.inv_2dthe * lt); break;
        case 3: aadd(&a.lav[q + R], a.inv_dthe * lt); aadd(&a.lav[q], -a.inv_dthe * lt); break;
        case 4: aadd(&a.lav[q - R], a.inv_dthe * lt); aadd(&a.lav[q], -a.inv_dthe * lt); break;
        default: aadd(&a.lav[q + R], a.inv_2dthe * lt); aadd(&a.lav[q - R], -a.inv_2dthe * lt); break;
    }
    switch(p_mode){
        case 1: aadd(&a.law[q], -3.0 * a.inv_2dphi * lp); aadd(&a.law[q + 1], 4.0 * a.inv_2dphi * lp);
                aadd(&a.law[q + 2], -a.inv_2dphi * lp); break;
        case 2: aadd(&a.law[q], 3.0 * a.inv_2dphi * lp); aadd(&a.law[q - 1], -4.0 * a.inv_2dphi * lp);
                aadd(&a.law[q - 2], a.inv_2dphi * lp); break;
        case 3: aadd(&a.law[q + 1], a.inv_dphi * lp); aadd(&a.law[q], -a.inv_dphi * lp); break;
        case 4: aadd(&a.law[q - 1], a.inv_dphi * lp); aadd(&a.law[q], -a.inv_dphi * lp); break;
        default: aadd(&a.law[q - k + ke], a.inv_2dphi * lp); aadd(&a.law[q - k + kw], -a.inv_2dphi * lp); break;
    }
}

// prepareProjectionBoundaries^T: aux_u zeroed at the radial walls, aux_v, aux_w 2-point Neumann
template<typename T>
__global__ void ka_prep_projection(T* __restrict__ lau, T* __restrict__ lav, T* __restrict__ law,
        int im, int jm, int km, int pitch, size_t plane, double c43, double c13){
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y + 1;
    if(k > km - 2) return;
    const size_t q = (size_t)j * pitch + k;
    const size_t t0 = q, t1 = q + plane, t2 = q + 2 * plane;
    const size_t b0 = q + (size_t)(im - 1) * plane, b1 = q + (size_t)(im - 2) * plane, b2 = q + (size_t)(im - 3) * plane;
    lau[t0] = 0.0;
    lau[b0] = 0.0;
    T* X[2] = { lav, law };
    for(int f = 0; f < 2; f++){
        T* A = X[f];
        A[t1] += c43 * A[t0]; A[t2] += -c13 * A[t0]; A[t0] = 0.0;
        A[b1] += c43 * A[b0]; A[b2] += -c13 * A[b0]; A[b0] = 0.0;
    }
}

// split <-> pitch layout for the adjoint of p (add into the target)
template<typename T>
__global__ void ka_rb_pack_add(const T* __restrict__ L, double* __restrict__ x0, double* __restrict__ x1,
        int jm, int km, int pitch, int hp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    double* x = ((i + j + k) & 1) ? x1 : x0;
    x[((size_t)i * jm + j) * hp + (k >> 1)] = L[((size_t)i * jm + j) * pitch + k];
}
template<typename T>
__global__ void ka_rb_unpack_add(T* __restrict__ L, const double* __restrict__ x0, const double* __restrict__ x1,
        int jm, int km, int pitch, int hp){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const double* x = ((i + j + k) & 1) ? x1 : x0;
    // the compact index of the seam (k = km - 1) is shared with nothing; every (i, j, k) maps once
    L[((size_t)i * jm + j) * pitch + k] = (T)x[((size_t)i * jm + j) * hp + (k >> 1)];
}

// the cells the RK4 step integrates (and overwrites aux_* in) lose their adjoint
template<typename T>
__global__ __launch_bounds__(256) void ka_zero_interior(T* __restrict__ a0, T* __restrict__ a1,
        T* __restrict__ a2, const uint8_t* __restrict__ flags, int im, int jm, int km, int pitch){
    const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if(k >= km) return;
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    const bool interior = i >= 1 && i <= im - 2 && j >= 3 && j <= jm - 4 && k <= km - 2 && !(flags[q] & FL_SELF);
    if(!interior) return;
    a0[q] = 0.0; a1[q] = 0.0; a2[q] = 0.0;
}

template<typename T>
__global__ void ka_zero(T* __restrict__ x, size_t n){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) x[q] = T(0);
}
template<typename T>
__global__ void ka_copy(T* __restrict__ d, const T* __restrict__ s, size_t n){
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if(q < n) d[q] = s[q];
}

}  // namespace atjup

#ifndef ATJUP_RBT32_PAIRS
#define ATJUP_RBT32_PAIRS 1
#endif

// =============================================================================================
// host side
// =============================================================================================
namespace atjup {

enum { AW_G = 0, AW_LK = 4, AW_LNP = 8, AW_Y = 12, AW_NN = 24, AW_AUX = 28, AW_AUXP = 31, AW_GBRL = 34,
       AW_SCR = 35, AW_PIN = 36, AW_N = 37 };
// the forward state a step starts from, as checkpoints and snapshots hold it; the FP32-refined
// form adds the low words of its paired fields t, u, v, w
static const int kStateFields[] = { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W, ATJUP_F_P_DYN,
                                    ATJUP_F_AUX_U, ATJUP_F_AUX_V, ATJUP_F_AUX_W, ATJUP_F_RHO_MIX };
constexpr int N_STATE_FIELDS = 9, N_PAIRED = 4;
static int n_state(const atjup_ctx* c){
    return N_STATE_FIELDS + (c->precision == ATJUP_PRECISION_FP32_REF ? N_PAIRED : 0);
}

struct AdjState {
    Buf lam[ATJUP_F_END];
    Buf work[AW_N];
    Buf lsplit, lsrc, mu, inmask, red, zs;
    Buf ld32, lr32, mu32;         // the refined form's rounds: the correction's adjoint, the round's
                                  // source adjoint and the transposed FP32 sweeps' scaled adjoints
    std::vector<Buf> ckpt;        // n_state fields per checkpoint
    std::vector<Buf> snap;        // n_state fields per step of the segment being reversed
    std::vector<Buf> post;        // N_POST fields per physics-block step of the whole run: the state
                                  // after its block (the pressure solve's result), kept from the forward
    std::vector<int> post_slot;   // step -> its slot in post, -1 for steps without the block
    // a captured transposed solve, replayed only on the buffers and geometry it was captured with
    // (its kernels' arguments are baked in); kind 0: FP64 sweeps, 1: the refined form's rounds
    struct SweepGraph { int kind; int n; double omega; int nr; hipGraphExec_t exec; const void* ptrs[12]; int dims[4]; };
    std::vector<SweepGraph> graphs;
    Buf brl_part;                 // per-level partial sums of the level-mean transpose
    Buf gchi;                     // dL/dchi of the drag region, [n3]
    Buf obst_part;                // per-block partial sums of dL/d(semi-axes)
    int nsteps = 0, every = 0, first_iter = 1;
    bool last_in_snap = false;    // snap holds the last segment's step states, as the forward left them
};

// what the physics block of a dynamics-profile iteration changes: p_dyn (solve and smoothing),
// aux_* (prepareProjectionBoundaries) and rho_mix
static const int kPostFields[] = { ATJUP_F_P_DYN, ATJUP_F_AUX_U, ATJUP_F_AUX_V, ATJUP_F_AUX_W, ATJUP_F_RHO_MIX };
constexpr int N_POST = 5;

static const int kLamFields[] = { ATJUP_F_T, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W, ATJUP_F_P_DYN, ATJUP_F_P_HYDRO,
                                  ATJUP_F_RHO_MIX, ATJUP_F_AUX_U, ATJUP_F_AUX_V, ATJUP_F_AUX_W };

static int adj_alloc(Buf& b, size_t bytes){
    if(hipMalloc(&b.p, bytes) != hipSuccess){ set_error("hipMalloc failed (adjoint)"); return ATJUP_ERR_NOMEM; }
    b.bytes = bytes;
    if(hipMemset(b.p, 0, bytes) != hipSuccess){ set_error("hipMemset failed"); return ATJUP_ERR_HIP; }
    return 0;
}
static void adj_release(Buf& b){ if(b.p) (void)hipFree(b.p); b.p = nullptr; b.bytes = 0; }

void adj_free(atjup_ctx* c){
    if(!c || !c->adj) return;
    AdjState* A = c->adj;
    for(auto& b : A->lam) adj_release(b);
    for(auto& b : A->work) adj_release(b);
    for(Buf* b : { &A->lsplit, &A->lsrc, &A->mu, &A->inmask, &A->red, &A->zs, &A->brl_part, &A->gchi, &A->obst_part,
                   &A->ld32, &A->lr32, &A->mu32 })
        adj_release(*b);
    for(auto& b : A->ckpt) adj_release(b);
    for(auto& b : A->snap) adj_release(b);
    for(auto& b : A->post) adj_release(b);
    for(auto& g : A->graphs) (void)hipGraphExecDestroy(g.exec);
    delete A;
    c->adj = nullptr;
}

// the adjoint fields, work fields and forward fields in the context's storage type T
template<typename T> static inline T* L(atjup_ctx* c, int f){ return (T*)c->adj->lam[f].p; }
template<typename T> static inline T* W(atjup_ctx* c, int w){ return (T*)c->adj->work[w].p; }
template<typename T> static inline T* Fd(atjup_ctx* c, int f){ return (T*)c->cur[f].p; }

static int zero_buf(atjup_ctx* c, Buf& b){
    const size_t n = b.bytes / sizeof(uint32_t);            // every adjoint buffer is a multiple of 4 bytes
    hipLaunchKernelGGL(ka_zero<uint32_t>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, (uint32_t*)b.p, n);
    return check_launch();
}
template<typename T>
static int copy_field(atjup_ctx* c, T* d, const T* s){
    hipLaunchKernelGGL(ka_copy<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream, d, s, c->n3);
    return check_launch();
}

// ---- the transposed pieces --------------------------------------------------------------------
template<typename T>
static int adj_bc(atjup_ctx* c){
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    T *lt = L<T>(c, ATJUP_F_T), *lu = L<T>(c, ATJUP_F_U), *lv = L<T>(c, ATJUP_F_V), *lw = L<T>(c, ATJUP_F_W);
    T* lp = L<T>(c, ATJUP_F_P_DYN);
    if(c->n_amb_rows > 0)
        hipLaunchKernelGGL(ka_t_ambient<T>, dim3(c->n_amb_rows), dim3(64), 0, c->stream, lt,
            (const uint8_t*)c->mask.p, (const int2*)c->amb_rows.p, c->n_amb_rows, c->jm, c->km, c->pitch);
    if(c->n_solids > 0)
        hipLaunchKernelGGL(ka_bc_solid<T>, dim3(nblk(c->n_solids, 64)), dim3(64), 0, c->stream, lu, lv, lw, lt, lp,
            (const SolidCell*)c->solids.p, c->n_solids, c->jm, c->pitch, c->plane, c43, c13);
    hipLaunchKernelGGL(ka_bc_phi<T>, dim3(nblk(c->jm, 64), c->im), dim3(64), 0, c->stream, lt, lu, lv, lw,
        c->jm, c->km, c->pitch);
    hipLaunchKernelGGL(ka_bc_theta<T>, dim3(nblk(c->km, 128), c->im), dim3(128), 0, c->stream, lt, lu, lv, lw,
        c->jm, c->km, c->pitch);
    double* zs = (double*)c->adj->zs.p;
    hipLaunchKernelGGL(ka_bc_zonal_sums<T>, dim3(nblk(c->jm, 64), 4), dim3(64), 0, c->stream, (const T*)lv,
        (const T*)lw, zs, c->im, c->jm, c->km, c->pitch, c->plane);
    hipLaunchKernelGGL(ka_bc_radius<T>, dim3(nblk(c->km, 128), c->jm), dim3(128), 0, c->stream, lt, lu, lv, lw,
        (const double*)zs, c->im, c->jm, c->km, c->pitch, c->plane, c43, c13);
    return check_launch();
}

template<typename T>
static int adj_shapiro(atjup_ctx* c, int f){
    // forward passes k, j, i (strength 1, order 2): transposed i, j, k through the scratch buffer
    T* a = L<T>(c, f);
    T* b = W<T>(c, AW_SCR);
    const double co = 0.25;
    dim3 g(nblk(c->km, 128), c->jm, c->im), blk(128);
    hipLaunchKernelGGL(ka_shapiro_i<T>, g, blk, 0, c->stream, (const T*)a, b, (const int*)c->itopo.p,
        c->im, c->jm, c->km, c->pitch, c->plane, co);
    hipLaunchKernelGGL(ka_shapiro_j<T>, g, blk, 0, c->stream, (const T*)b, a, (const int*)c->itopo.p,
        c->im, c->jm, c->km, c->pitch, co);
    hipLaunchKernelGGL(ka_shapiro_k<T>, g, blk, 0, c->stream, (const T*)a, b, (const int*)c->itopo.p,
        c->im, c->jm, c->km, c->pitch, co);
    return copy_field<T>(c, a, b);
}

template<typename T>
static int adj_mixture(atjup_ctx* c){
    hipLaunchKernelGGL(ka_mixture<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream,
        (const T*)Fd<T>(c, ATJUP_F_T), (const T*)Fd<T>(c, ATJUP_F_P_STAT), (const T*)Fd<T>(c, ATJUP_F_P_DYN),
        L<T>(c, ATJUP_F_RHO_MIX), L<T>(c, ATJUP_F_T), L<T>(c, ATJUP_F_P_DYN), c->n3, c->c.t_ref, c->kn.rho_pdyn,
        c->c.r_mix * c->c.u_0 * c->c.u_0 * 1.0e-5, c->c.R_mix);
    return check_launch();
}

// RK4^T at the context's state (the step's n level, p_dyn, rho_mix): records the stage inputs, then
// runs the four transposed stages; the n level's adjoint replaces the step result's, p_dyn and
// rho_mix gather theirs, the hydrostatic and level-mean adjoints are spread back onto v, w and t,
// and aux_* lose their adjoint in the cells the step overwrote.
template<typename T>
static int adj_rk4(atjup_ctx* c){
    AdjState* A = c->adj;
    RkBuffers rb{};
    for(int s = 0; s < 3; s++) for(int f = 0; f < 4; f++) rb.y[s][f] = (double*)W<T>(c, AW_Y + 4 * s + f);
    for(int f = 0; f < 4; f++) rb.nn[f] = (double*)W<T>(c, AW_NN + f);
    for(int f = 0; f < 3; f++) rb.aux[f] = (double*)W<T>(c, AW_AUX + f);
    if(int rc = launch_rk4_dry_record(c, rb)) return rc;
    const double dt = c->c.dt;
    const atjup_consts& C = c->c;
    RhsAdjArgs<T> a{};
    a.pd = Fd<T>(c, ATJUP_F_P_DYN); a.ph = Fd<T>(c, ATJUP_F_P_HYDRO); a.ps = Fd<T>(c, ATJUP_F_P_STAT);
    a.rho = Fd<T>(c, ATJUP_F_RHO_MIX); a.wn = Fd<T>(c, ATJUP_F_WALL_NUE);
    a.lpd = L<T>(c, ATJUP_F_P_DYN); a.lph = L<T>(c, ATJUP_F_P_HYDRO); a.lrho = L<T>(c, ATJUP_F_RHO_MIX);
    a.gbrl = W<T>(c, AW_GBRL);
    a.flags = (const uint8_t*)c->flags.p; a.geo = (const GeoIJ<geo_t<T>>*)c->geo.p;
    a.im = c->im; a.jm = c->jm; a.km = c->km; a.pitch = c->pitch; a.plane = c->plane;
    a.nkb = (c->km + 63) / 64; a.njb = (c->jm + 3) / 4; a.nb = a.nkb * a.njb * c->im;
    a.inv_2dr = 1.0 / (2.0 * c->g.dr); a.inv_2dthe = 1.0 / (2.0 * c->g.dthe); a.inv_2dphi = 1.0 / (2.0 * c->g.dphi);
    a.inv_dr2 = 1.0 / (c->g.dr * c->g.dr); a.inv_dthe2 = 1.0 / (c->g.dthe * c->g.dthe);
    a.inv_dphi2 = 1.0 / (c->g.dphi * c->g.dphi);
    a.cep = C.u_0 * C.u_0 / (C.cp_mix * C.t_ref);
    a.to_nd = 1.0 / (C.r_mix * C.u_0 * C.u_0 * 1.0e-5);
    a.r_mix = C.r_mix; a.c_m2o = 1.0 * -2.0 * C.omega; a.c_p2o = 1.0 * +2.0 * C.omega;
    a.Coriolis = C.Coriolis; a.cbuoy = -1.0 * 1.0 * 1.0 * C.buoyancy; a.g_ = C.g; a.rR = C.r_mix * C.R_mix;
    a.t_ref = C.t_ref; a.kt = c->kn.t_diff_scale / (C.re * C.pr) + 0.0 / 0.9; a.visc0 = 1.0 / C.re + 0.0;
    a.chi = c->obst_on ? (const double*)c->chi.p : nullptr;
    a.inv_tau = c->obst_on ? 1.0 / c->obst.tau : 0.0;
    a.gchi = (double*)A->gchi.p;
    for(int f = 0; f < 4; f++){
        a.lk[f] = W<T>(c, AW_LK + f);
        a.g[f] = W<T>(c, AW_G + f);
        if(int rc = copy_field<T>(c, W<T>(c, AW_LNP + f), L<T>(c, f))) return rc;      // lambda of the step result
    }
    dim3 gc(nblk(c->km, 128), c->jm, c->im), b128(128);
    auto combine = [&](double wnp, double cg, int have_g, int write_lk){
        hipLaunchKernelGGL(ka_rk_combine<T>, gc, b128, 0, c->stream,
            W<T>(c, AW_LK), W<T>(c, AW_LK + 1), W<T>(c, AW_LK + 2), W<T>(c, AW_LK + 3),
            W<T>(c, AW_G), W<T>(c, AW_G + 1), W<T>(c, AW_G + 2), W<T>(c, AW_G + 3),
            (const T*)W<T>(c, AW_LNP), (const T*)W<T>(c, AW_LNP + 1), (const T*)W<T>(c, AW_LNP + 2),
            (const T*)W<T>(c, AW_LNP + 3), L<T>(c, 0), L<T>(c, 1), L<T>(c, 2), L<T>(c, 3),
            (const uint8_t*)c->flags.p, c->im, c->jm, c->km, c->pitch, wnp, cg, have_g, write_lk);
        return check_launch();
    };
    const double wts[4] = { dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0 };
    const double cin[4] = { 0.5 * dt, 0.5 * dt, dt, 0.0 };
    for(int s = 3; s >= 0; s--){
        // lk_s = w_s lambda_N' + c_s lambda_{Y_{s+1}}; lambda_N += lambda_{Y_{s+1}}
        if(int rc = combine(wts[s], cin[s], s < 3 ? 1 : 0, 1)) return rc;
        for(int f = 0; f < 4; f++) a.y[f] = (s == 0) ? Fd<T>(c, f) : W<T>(c, AW_Y + 4 * (s - 1) + f);
        for(int f = 0; f < 3; f++) a.la[f] = (s == 3) ? L<T>(c, ATJUP_F_AUX_U + f) : nullptr;
        hipLaunchKernelGGL(ka_rhs_dry<T>, dim3(a.nb), dim3(64, 4), 0, c->stream, a, a.geo);
        if(int rc = check_launch()) return rc;
    }
    if(int rc = combine(0.0, 0.0, 1, 0)) return rc;             // lambda_N += lambda_{Y_0}
    // the level means and the hydrostatic integral, formed once from the n level
    const double* tabs = (const double*)c->tabs.p;
    hipLaunchKernelGGL(ka_brl_reduce<T>, dim3(ABRL_NCH, c->im), dim3(ABRL_BLOCK), 0, c->stream,
        (const T*)Fd<T>(c, ATJUP_F_T), (const T*)Fd<T>(c, ATJUP_F_P_STAT), (const uint8_t*)c->flags.p,
        tabs + c->im, (const T*)W<T>(c, AW_GBRL), (double*)A->brl_part.p, c->jm, c->km, c->pitch, C.g,
        C.r_mix * C.R_mix, C.t_ref);
    hipLaunchKernelGGL(ka_brl_final, dim3(nblk(c->im, 64)), dim3(64), 0, c->stream, (const double*)A->brl_part.p,
        (double*)A->red.p, c->im);
    hipLaunchKernelGGL(ka_brl_spread<T>, gc, b128, 0, c->stream, (const T*)Fd<T>(c, ATJUP_F_T),
        (const T*)Fd<T>(c, ATJUP_F_P_STAT), (const uint8_t*)c->flags.p, tabs + c->im, (const double*)A->red.p,
        W<T>(c, AW_GBRL), L<T>(c, ATJUP_F_T), c->jm, c->km, c->pitch, C.g, C.r_mix * C.R_mix, C.t_ref);
    hipLaunchKernelGGL(ka_hydrostatic<T>, dim3(nblk(c->km, 128), c->jm), b128, 0, c->stream,
        (const T*)Fd<T>(c, ATJUP_F_V), (const T*)Fd<T>(c, ATJUP_F_W), (const int2*)c->colr.p,
        tabs, tabs + c->im + c->jm, L<T>(c, ATJUP_F_P_HYDRO), L<T>(c, ATJUP_F_V), L<T>(c, ATJUP_F_W),
        c->im, c->jm, c->km, c->pitch, c->plane, c->g.dr);
    // aux_* were overwritten in the integrated cells: their earlier values carry no adjoint there
    hipLaunchKernelGGL(ka_zero_interior<T>, gc, b128, 0, c->stream, L<T>(c, ATJUP_F_AUX_U), L<T>(c, ATJUP_F_AUX_V),
        L<T>(c, ATJUP_F_AUX_W), (const uint8_t*)c->flags.p, c->im, c->jm, c->km, c->pitch);
    return check_launch();
}

// PressureSolver::run^T at the context's state (p_dyn and aux_* before the solve). The FP64 form
// transposes its sweeps; the FP32-refined form transposes what it runs: the first sweep in FP64,
// then rounds of FP32 sweeps on the correction inside the FP64 refinement (the rounds' FP32
// sweeps run transposed in FP32, the residuals and corrections transposed in FP64).
template<typename T>
static int adj_pressure(atjup_ctx* c, int sweeps, double omega){
    AdjState* A = c->adj;
    const double c43 = 4.0 / 3.0, c13 = 1.0 / 3.0;
    // the aux_* after prepareProjectionBoundaries, which the clamp test of the source reads (the
    // solve changes nothing else the transpose depends on: the sweeps are linear); p_dyn and aux_*
    // of the context are restored afterwards
    if(int rc = copy_field<T>(c, W<T>(c, AW_PIN), Fd<T>(c, ATJUP_F_P_DYN))) return rc;
    for(int f = 0; f < 3; f++)
        if(int rc = copy_field<T>(c, W<T>(c, AW_AUXP + f), Fd<T>(c, ATJUP_F_AUX_U + f))) return rc;
    if(int rc = launch_prep_projection(c)) return rc;
    const int hp = c->rb_hp;
    const size_t hn = (size_t)c->im * c->jm * hp;
    const int nm = (c->km - 1) / 2;
    double* l0 = (double*)A->lsplit.p;
    double* l1 = l0 + hn;
    double* s0 = (double*)A->lsrc.p;
    double* s1 = s0 + hn;
    const uint8_t* f0 = (const uint8_t*)c->rb_flags.p;
    const uint8_t* f1 = f0 + hn;
    const PoisIJ<double>* pw = (const PoisIJ<double>*)c->pois.p;
    if(int rc = zero_buf(c, A->lsrc)) return rc;
    dim3 gfull(nblk(c->km, 128), c->jm, c->im), b128(128);
    hipLaunchKernelGGL(ka_rb_pack_add<T>, gfull, b128, 0, c->stream, (const T*)L<T>(c, ATJUP_F_P_DYN), l0, l1,
        c->jm, c->km, c->pitch, hp);
    dim3 gseam(nblk(c->jm, 64), c->im), gpj(nblk(nm + 1, 128), c->im, 2), gpi(nblk(nm + 1, 128), c->jm, 2);
    dim3 gt(nblk(nm, RBT_BM), nblk(c->jm, RBT_BJ), c->im), bt(RBT_BM, RBT_BJ);
    dim3 gt2(nblk(nm, RBT_BM), nblk(c->jm, RBT_BJ), 2 * c->im);
    dim3 gcol(nblk(nm, 128), c->jm, 2 * c->im);
    uint8_t* in0 = (uint8_t*)A->inmask.p;
    uint8_t* in1 = in0 + hn;
    // two cells per thread, flattened over (row, pair), where the ring is even and row-aligned
    // pairs would leave more than 10% of the lanes idle; else one cell per thread. Measured on
    // MI355X over the misfit gradients of the bench: 41 x 181 x 361 (90 pairs per row) 81 ms of
    // transposed sweeps flattened against 84 ms one cell; 81 x 361 x 721 (180 pairs) 1478 ms
    // one cell against 1592 ms row-aligned pairs and 1729 ms flattened.
    const int wf_pair = (nm / 2 + RBT_BM - 1) / RBT_BM;
    const bool pairs64 = nm % 2 == 0 && (double)(nm / 2) / (RBT_BM * wf_pair) < 0.9;
    const int ntb = (c->jm * (nm / 2) + 255) / 256;
    // n sweeps transposed on the split adjoint la (its mu and source adjoint beside it), in the
    // sweeps' own precision; the forward's first sweep reads the stored planes unless all of them
    // form the planes (the refined form's correction sweeps)
    auto tsweeps = [&](auto* la0, auto* mua0, auto* lsa0, int n, bool all_ghost){
        using S = typename std::remove_pointer<decltype(la0)>::type;
        // FP32: pairs wherever the ring is even (ATJUP_RBT32_PAIRS 0: the FP64 rule)
        const bool pairs = pairs64 || (ATJUP_RBT32_PAIRS && std::is_same<S, float>::value && nm % 2 == 0);
        S* la1 = la0 + hn;
        S* mua1 = mua0 + hn;
        S* lsa1 = lsa0 + hn;
        hipLaunchKernelGGL(ka_rbT_mu<S>, gt, bt, 0, c->stream, (const S*)la1, mua1, f1, pw, c->im, c->jm, hp, nm, (S)omega);
        for(int s = n - 1; s >= 0; s--){
            const int ghost = (all_ghost || s > 0) ? 1 : 0;
            for(int colour = 0; colour < 2; colour++){
                S* lc = colour ? la1 : la0;
                S* mc = colour ? mua1 : mua0;
                const S* mo = colour ? mua0 : mua1;
                S* sc = colour ? lsa1 : lsa0;
                const uint8_t* fc = colour ? f1 : f0;
                const uint8_t* fo = colour ? f0 : f1;
                const uint8_t* ic = colour ? in1 : in0;
                if(pairs)
                    hipLaunchKernelGGL(ka_rbT_pass2<S>, dim3(ntb * c->im), dim3(256), 0, c->stream, lc, mc, mo, sc, fc, fo,
                        ic, pw, c->im, c->jm, hp, nm, ntb, colour, ghost, (S)omega, (S)c43, (S)c13);
                else
                    hipLaunchKernelGGL(ka_rbT_pass<S>, gt, bt, 0, c->stream, lc, mc, mo, sc, fc, fo,
                        ic, pw, c->im, c->jm, hp, nm, colour, ghost, (S)omega, (S)c43, (S)c13);
            }
        }
    };
    auto planes_T = [&](){      // k_rb_planes_i then k_rb_planes_j, transposed
        hipLaunchKernelGGL(ka_rb_planes_j, gpj, b128, 0, c->stream, l0, l1, c->im, c->jm, hp, nm, c43, c13);
        hipLaunchKernelGGL(ka_rb_planes_i, gpi, b128, 0, c->stream, l0, l1, c->im, c->jm, hp, nm, c43, c13);
    };
    const bool refined = std::is_same<T, float>::value && sweeps > 1;
    std::vector<int> rounds;                      // the refined form's correction sweeps per round
    if(refined)
        for(int left = sweeps - 1; left > 0;){ const int n = std::min(left, c->refine_sweeps); rounds.push_back(n); left -= n; }
    float* ld0 = (float*)A->ld32.p;
    float* lr0 = (float*)A->lr32.p;
    float* m32 = (float*)A->mu32.p;
    if(refined && (!ld0 || !lr0 || !m32)){ set_error("adjoint: the refined form's buffers are missing"); return ATJUP_ERR_STATE; }
    auto launch_all = [&](){
        hipLaunchKernelGGL(ka_rbT_inmask, gt2, bt, 0, c->stream, f0, f1, in0, in1, c->im, c->jm, hp, nm);
        // the planes and the seam written after the last sweep
        hipLaunchKernelGGL(ka_rb_seam, gseam, dim3(64), 0, c->stream, l0, l1, c->im, c->jm, hp, nm);
        planes_T();
        if(!refined){
            tsweeps(l0, (double*)A->mu.p, s0, sweeps, false);
            return;
        }
        for(int r = (int)rounds.size() - 1; r >= 0; r--){
            hipLaunchKernelGGL(ka_rb_addT, gcol, b128, 0, c->stream, (const double*)l0, (const double*)l1,
                ld0, ld0 + hn, lr0, lr0 + hn, c->im, c->jm, hp, nm);
            tsweeps(ld0, m32, lr0, rounds[r], true);
            hipLaunchKernelGGL(ka_rb_resT, gcol, b128, 0, c->stream, l0, l1, s0, s1, (const float*)lr0,
                (const float*)(lr0 + hn), f0, f1, (const uint8_t*)in0, (const uint8_t*)in1, pw, c->im, c->jm, hp, nm);
            planes_T();
        }
        tsweeps(l0, (double*)A->mu.p, s0, 1, false);          // the first sweep, FP64
    };
    // the transposed solve goes to the device as one graph, captured on first use per kind, sweep
    // count, relaxation factor and round length, and replayed only on the buffers it was captured with
    const int kind = refined ? 1 : 0;
    const void* ptrs[12] = { l0, s0, A->mu.p, in0, f0, pw, ld0, lr0, m32, nullptr, nullptr, nullptr };
    const int dims[4] = { c->im, c->jm, hp, nm };
    const int nr = refined ? c->refine_sweeps : 0;
    hipGraphExec_t exec = nullptr;
    AdjState::SweepGraph* slot = nullptr;
    for(auto& g : A->graphs) if(g.kind == kind && g.n == sweeps && g.omega == omega && g.nr == nr){ slot = &g; break; }
    if(slot && std::equal(ptrs, ptrs + 12, slot->ptrs) && std::equal(dims, dims + 4, slot->dims)) exec = slot->exec;
    if(!exec){
        hipGraph_t graph = nullptr;
        ATJUP_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        launch_all();
        const hipError_t le = hipGetLastError();
        ATJUP_HIP_CHECK(hipStreamEndCapture(c->stream, &graph));
        if(le != hipSuccess){ (void)hipGraphDestroy(graph); set_error(std::string("adjoint sweep capture: ") + hipGetErrorString(le)); return ATJUP_ERR_HIP; }
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if(ie != hipSuccess){ set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie)); return ATJUP_ERR_HIP; }
        AdjState::SweepGraph g{ kind, sweeps, omega, nr, exec, {}, {} };
        std::copy(ptrs, ptrs + 12, g.ptrs);
        std::copy(dims, dims + 4, g.dims);
        if(slot){
            ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
            (void)hipGraphExecDestroy(slot->exec);
            *slot = g;
        } else {
            A->graphs.push_back(g);
        }
    }
    hipEvent_t t_sw = timer_begin(c);
    ATJUP_HIP_CHECK(hipGraphLaunch(exec, c->stream));
    timer_end(c, ATJUP_T_ADJ_SWEEPS, t_sw, 2 * sweeps + 5);
    hipLaunchKernelGGL(ka_rb_unpack_add<T>, gfull, b128, 0, c->stream, L<T>(c, ATJUP_F_P_DYN), (const double*)l0,
        (const double*)l1, c->jm, c->km, c->pitch, hp);
    // the source and prepareProjectionBoundaries
    SrcAdjArgs<T> sa{};
    sa.au = Fd<T>(c, ATJUP_F_AUX_U); sa.av = Fd<T>(c, ATJUP_F_AUX_V); sa.aw = Fd<T>(c, ATJUP_F_AUX_W);
    sa.ls0 = s0; sa.ls1 = s1;
    sa.lau = L<T>(c, ATJUP_F_AUX_U); sa.lav = L<T>(c, ATJUP_F_AUX_V); sa.law = L<T>(c, ATJUP_F_AUX_W);
    sa.flags = (const uint8_t*)c->flags.p; sa.msk = (const uint8_t*)c->mask.p; sa.pw = pw;
    sa.im = c->im; sa.jm = c->jm; sa.km = c->km; sa.pitch = c->pitch; sa.hp = hp; sa.plane = c->plane;
    sa.inv_2dr = 1.0 / (2.0 * c->g.dr); sa.inv_2dthe = 1.0 / (2.0 * c->g.dthe); sa.inv_2dphi = 1.0 / (2.0 * c->g.dphi);
    sa.inv_dthe = 1.0 / c->g.dthe; sa.inv_dphi = 1.0 / c->g.dphi; sa.cap = c->kn.pdyn_cap;
    hipLaunchKernelGGL(ka_poisson_source<T>, dim3(nblk(c->km - 1, 128), c->jm - 2, c->im - 2), b128, 0, c->stream, sa);
    hipLaunchKernelGGL(ka_prep_projection<T>, dim3(nblk(c->km - 1, 128), c->jm - 2), b128, 0, c->stream,
        L<T>(c, ATJUP_F_AUX_U), L<T>(c, ATJUP_F_AUX_V), L<T>(c, ATJUP_F_AUX_W), c->im, c->jm, c->km, c->pitch, c->plane, c43, c13);
    // restore the context's state before the solve
    if(int rc = copy_field<T>(c, Fd<T>(c, ATJUP_F_P_DYN), W<T>(c, AW_PIN))) return rc;
    for(int f = 0; f < 3; f++)
        if(int rc = copy_field<T>(c, Fd<T>(c, ATJUP_F_AUX_U + f), W<T>(c, AW_AUXP + f))) return rc;
    return check_launch();
}

// one loop iteration transposed. The context holds the state the iteration starts from (t, u, v,
// w, p_dyn, aux_*, rho_mix); the forward physics block is rerun to reach the states its
// transposes and the RK4's linearise about. On return the adjoint fields have been carried from
// the end of the iteration to its start; the context's forward fields are left as the recording
// left them (the n level and aux_* at the iteration's start, p_dyn and rho_mix after the block).
template<typename T>
static int adj_step(atjup_ctx* c, int iter, int first, const Buf* post = nullptr){
    const bool phys = (iter % 2 == 0) || first;
    const int sweeps = std::max(1, c->kn.press_sweeps);
    if(phys){
        if(int rc = copy_field<T>(c, W<T>(c, AW_PIN), Fd<T>(c, ATJUP_F_P_DYN))) return rc;
        for(int f = 0; f < 3; f++)
            if(int rc = copy_field<T>(c, W<T>(c, AW_AUXP + f), Fd<T>(c, ATJUP_F_AUX_U + f))) return rc;
        if(post){
            // the state after the block, as the recomputation recorded it
            for(int s = 0; s < N_POST; s++)
                if(int rc = copy_field<T>(c, Fd<T>(c, kPostFields[s]), (const T*)post[s].p)) return rc;
        } else {
            if(int rc = launch_pressure(c, sweeps, c->kn.press_omega)) return rc;
            if(c->kn.damp_pdyn)
                if(int rc = launch_shapiro(c, ATJUP_F_P_DYN, 2, 1.0, 1, 7)) return rc;
            if(int rc = launch_mixture_density(c)) return rc;
        }
    }
    if(int rc = adj_bc<T>(c)) return rc;
    if(int rc = adj_rk4<T>(c)) return rc;
    if(phys){
        if(int rc = adj_mixture<T>(c)) return rc;
        if(c->kn.damp_pdyn)
            if(int rc = adj_shapiro<T>(c, ATJUP_F_P_DYN)) return rc;
        // back to the state before the solve for its transpose (adj_pressure keeps its own copy)
        T* pin = W<T>(c, AW_PIN);
        if(int rc = copy_field<T>(c, Fd<T>(c, ATJUP_F_P_DYN), pin)) return rc;
        for(int f = 0; f < 3; f++)
            if(int rc = copy_field<T>(c, Fd<T>(c, ATJUP_F_AUX_U + f), W<T>(c, AW_AUXP + f))) return rc;
        if(int rc = adj_pressure<T>(c, sweeps, c->kn.press_omega)) return rc;
    }
    return 0;
}

}  // namespace atjup

using namespace atjup;

namespace {

int need_adj(atjup_ctx* c){
    if(!c){ set_error("null context"); return ATJUP_ERR_ARG; }
    if(!c->adj){ set_error("call atjup_hip_adj_init first"); return ATJUP_ERR_STATE; }
    return 0;
}

bool fp32(const atjup_ctx* c){ return c->precision == ATJUP_PRECISION_FP32_REF; }

// the forward state a step starts from: the nine fields and, in the refined form, the low words of
// the paired fields (all of one field's size: n3 doubles, or n3 floats in the refined form)
template<typename T>
int state_copy(atjup_ctx* c, Buf* ckpt, bool save){
    for(int s = 0; s < N_STATE_FIELDS; s++){
        T* f = Fd<T>(c, kStateFields[s]);
        T* b = (T*)ckpt[s].p;
        if(int rc = save ? copy_field<T>(c, b, f) : copy_field<T>(c, f, b)) return rc;
    }
    for(int s = N_STATE_FIELDS; s < n_state(c); s++){
        float* f = (float*)c->lo[ATJUP_F_T + (s - N_STATE_FIELDS)].p;
        float* b = (float*)ckpt[s].p;
        if(!f){ set_error("adjoint: a paired field has no low word"); return ATJUP_ERR_STATE; }
        if(int rc = save ? copy_field<float>(c, b, f) : copy_field<float>(c, f, b)) return rc;
    }
    return 0;
}
int save_state(atjup_ctx* c, Buf* dst){
    return fp32(c) ? state_copy<float>(c, dst, true) : state_copy<double>(c, dst, true);
}
int load_state(atjup_ctx* c, const Buf* src){
    return fp32(c) ? state_copy<float>(c, const_cast<Buf*>(src), false) : state_copy<double>(c, const_cast<Buf*>(src), false);
}
int post_copy(atjup_ctx* c, Buf* post, bool save){
    for(int s = 0; s < N_POST; s++){
        int rc;
        if(fp32(c)){
            float* f = Fd<float>(c, kPostFields[s]);
            rc = save ? copy_field<float>(c, (float*)post[s].p, f) : copy_field<float>(c, f, (const float*)post[s].p);
        } else {
            double* f = Fd<double>(c, kPostFields[s]);
            rc = save ? copy_field<double>(c, (double*)post[s].p, f) : copy_field<double>(c, f, (const double*)post[s].p);
        }
        if(rc) return rc;
    }
    return 0;
}

// host (i, j, k) doubles <-> a device field of the context's storage type in the pitch layout
int host_to_pitch(atjup_ctx* c, void* dev, const double* src){
    const size_t esz = fp32(c) ? sizeof(float) : sizeof(double);
    std::vector<double> h(c->n3, 0.0);
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++)
            std::memcpy(h.data() + c->idx(i, j, 0), src + ((size_t)i * c->jm + j) * c->km, c->km * sizeof(double));
    std::vector<float> hf;
    const void* up = h.data();
    if(esz == sizeof(float)){ hf.assign(h.begin(), h.end()); up = hf.data(); }
    ATJUP_HIP_CHECK(hipMemcpyAsync(dev, up, c->n3 * esz, hipMemcpyHostToDevice, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}
int pitch_to_host(atjup_ctx* c, double* dst, const void* dev){
    const size_t esz = fp32(c) ? sizeof(float) : sizeof(double);
    std::vector<double> h(c->n3);
    if(esz == sizeof(float)){
        std::vector<float> hf(c->n3);
        ATJUP_HIP_CHECK(hipMemcpyAsync(hf.data(), dev, c->n3 * esz, hipMemcpyDeviceToHost, c->stream));
        ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
        for(size_t q = 0; q < c->n3; q++) h[q] = (double)hf[q];
    } else {
        ATJUP_HIP_CHECK(hipMemcpyAsync(h.data(), dev, c->n3 * esz, hipMemcpyDeviceToHost, c->stream));
        ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    for(int i = 0; i < c->im; i++)
        for(int j = 0; j < c->jm; j++)
            std::memcpy(dst + ((size_t)i * c->jm + j) * c->km, h.data() + c->idx(i, j, 0), c->km * sizeof(double));
    return 0;
}

// the transposed pieces in the context's precision
int adj_bc_any(atjup_ctx* c){ return fp32(c) ? adj_bc<float>(c) : adj_bc<double>(c); }
int adj_rk4_any(atjup_ctx* c){ return fp32(c) ? adj_rk4<float>(c) : adj_rk4<double>(c); }
int adj_mixture_any(atjup_ctx* c){ return fp32(c) ? adj_mixture<float>(c) : adj_mixture<double>(c); }
int adj_shapiro_any(atjup_ctx* c, int f){ return fp32(c) ? adj_shapiro<float>(c, f) : adj_shapiro<double>(c, f); }
int adj_pressure_any(atjup_ctx* c, int sweeps, double omega){
    return fp32(c) ? adj_pressure<float>(c, sweeps, omega) : adj_pressure<double>(c, sweeps, omega);
}
int adj_step_any(atjup_ctx* c, int iter, int first, const Buf* post = nullptr){
    return fp32(c) ? adj_step<float>(c, iter, first, post) : adj_step<double>(c, iter, first, post);
}

}  // namespace

extern "C" {

int atjup_hip_adj_init(atjup_ctx* c){
    if(!c){ set_error("null context"); return ATJUP_ERR_ARG; }
    if((c->precision != ATJUP_PRECISION_FP64 && c->precision != ATJUP_PRECISION_FP32_REF) || !c->kn.dry){
        set_error("the adjoint covers the dynamics profile in the FP64 and FP32-refined forms"); return ATJUP_ERR_PROFILE;
    }
    if(c->adj) return 0;
    c->adj = new AdjState();
    AdjState* A = c->adj;
    const size_t fb = c->n3 * (fp32(c) ? sizeof(float) : sizeof(double));
    int rc = 0;
    for(int f : kLamFields) if(!rc) rc = adj_alloc(A->lam[f], fb);
    for(int w = 0; w < AW_N && !rc; w++) rc = adj_alloc(A->work[w], fb);
    const size_t hn = (size_t)c->im * c->jm * c->rb_hp;
    if(!rc) rc = adj_alloc(A->lsplit, 2 * hn * sizeof(double));
    if(!rc) rc = adj_alloc(A->lsrc, 2 * hn * sizeof(double));
    if(!rc) rc = adj_alloc(A->mu, 2 * hn * sizeof(double));
    if(!rc) rc = adj_alloc(A->inmask, 2 * hn);
    if(!rc) rc = adj_alloc(A->red, 2 * (size_t)c->im * sizeof(double));
    if(!rc) rc = adj_alloc(A->brl_part, 2 * (size_t)c->im * ABRL_NCH * sizeof(double));
    if(!rc) rc = adj_alloc(A->gchi, c->n3 * sizeof(double));
    if(!rc) rc = adj_alloc(A->obst_part, 3 * (size_t)OBST_BLOCKS * sizeof(double));
    if(!rc) rc = adj_alloc(A->zs, 4 * (size_t)c->jm * sizeof(double));
    if(fp32(c)){
        if(!rc) rc = adj_alloc(A->ld32, 2 * hn * sizeof(float));
        if(!rc) rc = adj_alloc(A->lr32, 2 * hn * sizeof(float));
        if(!rc) rc = adj_alloc(A->mu32, 2 * hn * sizeof(float));
    }
    if(rc){ adj_free(c); return rc; }
    return 0;
}

int atjup_hip_adj_zero(atjup_ctx* c){
    if(int rc = need_adj(c)) return rc;
    for(int f : kLamFields) if(int rc = zero_buf(c, c->adj->lam[f])) return rc;
    if(int rc = zero_buf(c, c->adj->work[AW_GBRL])) return rc;
    for(int f = 0; f < 4; f++) if(int rc = zero_buf(c, c->adj->work[AW_G + f])) return rc;
    return zero_buf(c, c->adj->gchi);
}

int atjup_hip_set_smooth_obstacle(atjup_ctx* c, const atjup_obstacle* p){
    if(!c){ set_error("null context"); return ATJUP_ERR_ARG; }
    if(!p){ c->obst_on = false; return 0; }
    if(!c->kn.dry){ set_error("the smooth drag region belongs to the dynamics profile"); return ATJUP_ERR_PROFILE; }
    const double v[5] = { p->a_phi_deg, p->b_the_deg, p->c_r_km, p->width, p->tau };
    for(double x : v)
        if(!(x > 0.0) || !std::isfinite(x)){ set_error("semi-axes, width and tau must be positive"); return ATJUP_ERR_ARG; }
    if(!std::isfinite(p->j0) || !std::isfinite(p->k0)){ set_error("centre"); return ATJUP_ERR_ARG; }
    if(!c->chi.p){
        const size_t bytes = c->n3 * sizeof(double);
        if(hipMalloc(&c->chi.p, bytes) != hipSuccess){ c->chi.p = nullptr; set_error("hipMalloc failed"); return ATJUP_ERR_NOMEM; }
        c->chi.bytes = bytes;
        ATJUP_HIP_CHECK(hipMemsetAsync(c->chi.p, 0, bytes, c->stream));
    }
    c->obst = *p;
    const ObstGeom o = obst_geom(c);
    hipLaunchKernelGGL(k_obstacle_chi, dim3(nblk(c->km, 256), c->jm, c->im), dim3(256), 0, c->stream,
                       (double*)c->chi.p, o);
    if(int rc = check_launch()) return rc;
    c->obst_on = true;
    return 0;
}

int atjup_hip_adj_obstacle_gradient(atjup_ctx* c, double* out){
    if(int rc = need_adj(c)) return rc;
    if(!out){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(!c->obst_on){ set_error("no smooth drag region is set (atjup_hip_set_smooth_obstacle)"); return ATJUP_ERR_STATE; }
    AdjState* A = c->adj;
    const ObstGeom o = obst_geom(c);
    hipLaunchKernelGGL(ka_obstacle_grad, dim3(OBST_BLOCKS), dim3(256), 0, c->stream, (const double*)A->gchi.p,
                       (double*)A->obst_part.p, o);
    if(int rc = check_launch()) return rc;
    std::vector<double> h(3 * (size_t)OBST_BLOCKS);
    ATJUP_HIP_CHECK(hipMemcpyAsync(h.data(), A->obst_part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    out[0] = out[1] = out[2] = 0.0;
    for(int b = 0; b < OBST_BLOCKS; b++)
        for(int p = 0; p < 3; p++) out[p] += h[(size_t)b * 3 + p];
    return 0;
}

int atjup_hip_adj_upload(atjup_ctx* c, int f, const double* src){
    if(int rc = need_adj(c)) return rc;
    if(f < 0 || f >= ATJUP_F_END || !c->adj->lam[f].p || !src){ set_error("adjoint field id"); return ATJUP_ERR_ARG; }
    return host_to_pitch(c, c->adj->lam[f].p, src);
}

int atjup_hip_adj_download(atjup_ctx* c, int f, double* dst){
    if(int rc = need_adj(c)) return rc;
    if(f < 0 || f >= ATJUP_F_END || !c->adj->lam[f].p || !dst){ set_error("adjoint field id"); return ATJUP_ERR_ARG; }
    return pitch_to_host(c, dst, c->adj->lam[f].p);
}

int atjup_hip_adj_op(atjup_ctx* c, int op, int iarg, double darg){
    if(int rc = need_adj(c)) return rc;
    if(!c->have_mask){ set_error("upload the obstacle mask first"); return ATJUP_ERR_STATE; }
    switch(op){
        case ATJUP_ADJ_OP_BC:           return adj_bc_any(c);
        case ATJUP_ADJ_OP_SHAPIRO_PDYN: return adj_shapiro_any(c, ATJUP_F_P_DYN);
        case ATJUP_ADJ_OP_MIXTURE:      return adj_mixture_any(c);
        case ATJUP_ADJ_OP_RK4:          return adj_rk4_any(c);
        case ATJUP_ADJ_OP_PRESSURE:
            if(!(darg > 0.0 && darg < 2.0)){ set_error("omega outside (0, 2)"); return ATJUP_ERR_ARG; }
            return adj_pressure_any(c, iarg > 0 ? iarg : std::max(1, c->kn.press_sweeps), darg);
        case ATJUP_ADJ_OP_STEP:         return adj_step_any(c, iarg, darg != 0.0 ? 1 : 0);
        default: set_error("adjoint operator"); return ATJUP_ERR_ARG;
    }
}

int atjup_hip_adj_forward_checkpointed(atjup_ctx* c, int nsteps, int every, int first_iter){
    if(int rc = need_adj(c)) return rc;
    if(nsteps < 1 || every < 1){ set_error("nsteps, every"); return ATJUP_ERR_ARG; }
    if(!c->have_mask){ set_error("upload the obstacle mask first (atjup_hip_upload_mask)"); return ATJUP_ERR_STATE; }
    AdjState* A = c->adj;
    A->nsteps = 0;
    A->last_in_snap = false;
    // Storage: the state every `every` steps (the checkpoints), the state after each physics block
    // (the pressure solve's result: the recomputation of a segment then runs no solve), and the
    // step states of the last segment (its reverse sweep then needs no recomputation). On the
    // 41 x 181 x 361 shell, 20 steps with checkpoints every 5, FP64: 36 + 55 + 45 fields of 21.6 MB
    // (the refined form: 52 + 55 + 65 fields of 11.4 MB). Buffers are kept between calls of the
    // same shape.
    const int nseg = (nsteps + every - 1) / every;
    const int s_last = (nseg - 1) * every;
    const int ns = n_state(c);
    A->post_slot.assign(nsteps, -1);
    int nphys = 0;
    for(int n = 0; n < nsteps; n++){
        const int iter = first_iter + n;
        if(iter % 2 == 0 || n == 0) A->post_slot[n] = nphys++;
    }
    const size_t fb = c->n3 * (fp32(c) ? sizeof(float) : sizeof(double));
    auto ensure = [&](std::vector<Buf>& v, size_t count) -> int {
        if(v.size() == count && (count == 0 || v[0].bytes == fb)) return 0;
        for(auto& b : v) adj_release(b);
        v.assign(count, Buf{});
        for(auto& b : v) if(int rc = adj_alloc(b, fb)) return rc;
        return 0;
    };
    if(int rc = ensure(A->ckpt, (size_t)nseg * ns)) return rc;
    if(int rc = ensure(A->snap, (size_t)every * ns)) return rc;
    if(int rc = ensure(A->post, (size_t)nphys * N_POST)) return rc;
    for(int n = 0; n < nsteps; n++){
        const int iter = first_iter + n, first = n == 0 ? 1 : 0;
        if(n % every == 0)
            if(int rc = save_state(c, &A->ckpt[(size_t)(n / every) * ns])) return rc;
        if(n >= s_last)
            if(int rc = save_state(c, &A->snap[(size_t)(n - s_last) * ns])) return rc;
        hipEvent_t t_step = timer_begin(c);
        if(int rc = step_block(c, iter, first)) return rc;
        if(A->post_slot[n] >= 0)
            if(int rc = post_copy(c, &A->post[(size_t)A->post_slot[n] * N_POST], true)) return rc;
        if(int rc = step_rest(c, iter)) return rc;
        timer_end(c, ATJUP_T_STEP, t_step, 0);
    }
    A->nsteps = nsteps; A->every = every; A->first_iter = first_iter;
    A->last_in_snap = true;
    return 0;
}

int atjup_hip_adj_backward(atjup_ctx* c, const double* lam_final, double* lam_init){
    if(int rc = need_adj(c)) return rc;
    AdjState* A = c->adj;
    if(A->nsteps < 1){ set_error("run atjup_hip_adj_forward_checkpointed first"); return ATJUP_ERR_STATE; }
    if(lam_final){
        if(int rc = atjup_hip_adj_zero(c)) return rc;
        for(int f = 0; f < 4; f++)
            if(int rc = host_to_pitch(c, c->adj->lam[f].p, lam_final + (size_t)f * c->im * c->jm * c->km)) return rc;
    }
    const int every = A->every, nsteps = A->nsteps;
    const int nseg = (nsteps + every - 1) / every;
    const int ns = n_state(c);
    for(int seg = nseg - 1; seg >= 0; seg--){
        const int s0 = seg * every, s1 = std::min(nsteps, s0 + every);
        if(!(seg == nseg - 1 && A->last_in_snap)){
            // recompute the segment from its checkpoint, keeping each step's start; the physics
            // blocks' results come from the forward's store, so no pressure solve runs here
            if(int rc = load_state(c, &A->ckpt[(size_t)seg * ns])) return rc;
            hipEvent_t t_rc = timer_begin(c);
            for(int n = s0; n < s1; n++){
                const int iter = A->first_iter + n;
                if(int rc = save_state(c, &A->snap[(size_t)(n - s0) * ns])) return rc;
                if(n + 1 == s1) break;
                if(A->post_slot[n] >= 0)
                    if(int rc = post_copy(c, &A->post[(size_t)A->post_slot[n] * N_POST], false)) return rc;
                if(int rc = step_rest(c, iter)) return rc;
            }
            timer_end(c, ATJUP_T_ADJ_RECOMPUTE, t_rc, 0);
        }
        A->last_in_snap = false;        // the snapshots now belong to this segment
        for(int n = s1 - 1; n >= s0; n--){
            const int iter = A->first_iter + n, first = n == 0 ? 1 : 0;
            if(int rc = load_state(c, &A->snap[(size_t)(n - s0) * ns])) return rc;
            const int slot = A->post_slot[n];
            hipEvent_t t_as = timer_begin(c);
            if(int rc = adj_step_any(c, iter, first, slot >= 0 ? &A->post[(size_t)slot * N_POST] : nullptr)) return rc;
            timer_end(c, ATJUP_T_ADJ_STEP, t_as, 0);
        }
    }
    if(lam_init)
        for(int f = 0; f < 4; f++)
            if(int rc = pitch_to_host(c, lam_init + (size_t)f * c->im * c->jm * c->km, c->adj->lam[f].p)) return rc;
    return 0;
}

}  // extern "C"
