// Reporting reductions on the device (planet/Reporting.h): searchMinMax_3D / searchMinMax_2D
// (:261-330), the continuity residual of steadyQuery (:372-398), and slices of a field for the
// writers. Each reduction is exact: candidates carry their value and their position in the
// reference's loop order, and the combine keeps the reference's winner on ties (the first cell
// for the searches, the last for the residual), so the result does not depend on the grid's
// split into workgroups.
#include "atjup_internal.h"
#include "atjup_device.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>

namespace atjup {
namespace {

constexpr int RP_BK = 256, RP_BLOCKS = 2048;

struct Cand { double v; long long o; };
struct Part { Cand mx, mn; };

// max: larger value, then the earlier position; min: smaller value, then the earlier position;
// last-max (the residual's >= scan): larger value, then the later position
__device__ __forceinline__ bool better_max(const Cand& a, const Cand& b){ return a.v > b.v || (a.v == b.v && a.o < b.o); }
__device__ __forceinline__ bool better_min(const Cand& a, const Cand& b){ return a.v < b.v || (a.v == b.v && a.o < b.o); }
__device__ __forceinline__ bool better_last(const Cand& a, const Cand& b){ return a.v > b.v || (a.v == b.v && a.o > b.o); }

__device__ __forceinline__ Cand shfl_cand(const Cand& c, int m){
    Cand r;
    r.v = __shfl_xor(c.v, m);
    r.o = __shfl_xor(c.o, m);
    return r;
}

// reduce (mx, mn) over the workgroup; thread 0 holds the result
template<int MODE>   // 0: max / min pair, 1: last-max only (in mx)
__device__ void block_reduce(Cand& mx, Cand& mn){
    __shared__ Cand s_mx[RP_BK / 64], s_mn[RP_BK / 64];
    for(int m = 32; m >= 1; m >>= 1){
        const Cand a = shfl_cand(mx, m);
        if(MODE == 0 ? better_max(a, mx) : better_last(a, mx)) mx = a;
        if(MODE == 0){ const Cand b = shfl_cand(mn, m); if(better_min(b, mn)) mn = b; }
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    if(lane == 0){ s_mx[w] = mx; s_mn[w] = mn; }
    __syncthreads();
    if(threadIdx.x == 0)
        for(int x = 1; x < (int)(blockDim.x / 64); x++){
            if(MODE == 0 ? better_max(s_mx[x], mx) : better_last(s_mx[x], mx)) mx = s_mx[x];
            if(MODE == 0 && better_min(s_mn[x], mn)) mn = s_mn[x];
        }
}

__device__ __forceinline__ Cand no_max(){ return Cand{ -INFINITY, LLONG_MAX }; }
__device__ __forceinline__ Cand no_min(){ return Cand{ INFINITY, LLONG_MAX }; }

// searchMinMax_3D's candidates: every cell, position (j * km + k) * im + i (the loops run j, k, i);
// a NaN never wins a comparison, so it is never a candidate
template<typename T>
__global__ __launch_bounds__(RP_BK) void k_minmax3d(const T* __restrict__ f, const float* __restrict__ lo,
        int im, int jm, int km, int pitch, Part* __restrict__ parts){
    Cand mx = no_max(), mn = no_min();
    const size_t plane = (size_t)jm * pitch, n = (size_t)im * plane;
    for(size_t q = (size_t)blockIdx.x * RP_BK + threadIdx.x; q < n; q += (size_t)gridDim.x * RP_BK){
        const int i = (int)(q / plane);
        const size_t r = q - (size_t)i * plane;
        const int j = (int)(r / pitch), k = (int)(r - (size_t)j * pitch);
        if(k >= km) continue;
        double v = (double)f[q];
        if(lo) v += (double)lo[q];
        if(v != v) continue;
        const Cand c{ v, ((long long)j * km + k) * im + i };
        if(better_max(c, mx)) mx = c;
        if(better_min(c, mn)) mn = c;
    }
    block_reduce<0>(mx, mn);
    if(threadIdx.x == 0) parts[blockIdx.x] = Part{ mx, mn };
}

// searchMinMax_2D's candidates: the first element (position -1, the loops' starting value) and the
// interior j in [1, jm-2], k in [1, km-2], position j * km + k
__global__ __launch_bounds__(RP_BK) void k_minmax2d(const double* __restrict__ f, int jm, int km, Part* __restrict__ parts){
    Cand mx = no_max(), mn = no_min();
    const int n = jm * km;
    for(int q = blockIdx.x * RP_BK + threadIdx.x; q < n; q += gridDim.x * RP_BK){
        const int j = q / km, k = q - j * km;
        const bool first = q == 0, inner = j >= 1 && j <= jm - 2 && k >= 1 && k <= km - 2;
        if(!first && !inner) continue;
        const double v = f[q];
        if(v != v) continue;
        const Cand c{ v, first ? -1LL : (long long)q };
        if(better_max(c, mx)) mx = c;
        if(better_min(c, mn)) mn = c;
    }
    block_reduce<0>(mx, mn);
    if(threadIdx.x == 0) parts[blockIdx.x] = Part{ mx, mn };
}

// steadyQuery's continuity residual (Reporting.h:372-398) over the interior, |residual| with the
// last cell of the i, j, k loops on ties (>=, starting from 0)
template<typename T>
__global__ __launch_bounds__(RP_BK) void k_continuity(const T* __restrict__ u, const T* __restrict__ v, const T* __restrict__ w,
        const float* __restrict__ lu, const float* __restrict__ lv, const float* __restrict__ lw,
        const double* __restrict__ tabs, int im, int jm, int km, int pitch, double dr, double dthe, double dphi,
        Part* __restrict__ parts){
    Cand best{ -1.0, -1 }, unused = no_min();
    const int ni = im - 2, nj = jm - 2, nk = km - 2;
    const size_t n = (size_t)ni * nj * nk;
    const size_t P = (size_t)jm * pitch;
    auto ld = [&](const T* X, const float* L, size_t q){ double x = (double)X[q]; if(L) x += (double)L[q]; return x; };
    for(size_t e = (size_t)blockIdx.x * RP_BK + threadIdx.x; e < n; e += (size_t)gridDim.x * RP_BK){
        const int k = (int)(e % nk) + 1;
        const size_t r = e / nk;
        const int j = (int)(r % nj) + 1, i = (int)(r / nj) + 1;
        const size_t q = ((size_t)i * jm + j) * pitch + k;
        const double sinthe = tabs[im + j], costhe = tabs[im + 2 * jm + j];
        const double rm_met = tabs[i];
        const double rmsinthe = rm_met * sinthe;
        const double dudr = (ld(u, lu, q + P) - ld(u, lu, q - P)) / (2. * dr);
        const double dvdthe = (ld(v, lv, q + pitch) - ld(v, lv, q - pitch)) / (2. * dthe);
        const double dwdphi = (ld(w, lw, q + 1) - ld(w, lw, q - 1)) / (2. * dphi);
        const double res = dudr + 2. * ld(u, lu, q) / rm_met + dvdthe / rm_met
            + costhe / rmsinthe * ld(v, lv, q) + dwdphi / rmsinthe;
        const double a = fabs(res);
        if(!(a >= 0.0)) continue;                          // NaN
        const Cand c{ a, ((long long)i * jm + j) * km + k };
        if(better_last(c, best)) best = c;
    }
    block_reduce<1>(best, unused);
    if(threadIdx.x == 0) parts[blockIdx.x] = Part{ best, unused };
}

// one workgroup over the partial results; a NaN first element wins both searches, as the
// reference's loops start from it and no comparison with it succeeds
template<int MODE>
__global__ __launch_bounds__(RP_BK) void k_reduce_parts(const Part* __restrict__ parts, int nparts,
        const double* __restrict__ first, Part* __restrict__ out){
    Cand mx = MODE == 0 ? no_max() : Cand{ -1.0, -1 }, mn = no_min();
    for(int x = threadIdx.x; x < nparts; x += RP_BK){
        if(MODE == 0 ? better_max(parts[x].mx, mx) : better_last(parts[x].mx, mx)) mx = parts[x].mx;
        if(MODE == 0 && better_min(parts[x].mn, mn)) mn = parts[x].mn;
    }
    block_reduce<MODE>(mx, mn);
    if(threadIdx.x == 0){
        if(MODE == 0 && first && !(*first == *first)){ mx = Cand{ *first, 0 }; mn = mx; }
        *out = Part{ mx, mn };
    }
}

// steadyQuery's per-field change (Reporting.h:399-410): |f - f_n| over every cell, the last cell of
// the i, j, k loops on ties (>= from 0), position (i * jm + j) * km + k; NaNs never taken
template<typename T>
__global__ __launch_bounds__(RP_BK) void k_steady(const T* __restrict__ f, const float* __restrict__ lo,
        const double* __restrict__ fn, int im, int jm, int km, int pitch, Part* __restrict__ parts){
    Cand best{ -1.0, -1 }, unused = no_min();
    const size_t plane = (size_t)jm * pitch, n = (size_t)im * plane;
    for(size_t q = (size_t)blockIdx.x * RP_BK + threadIdx.x; q < n; q += (size_t)gridDim.x * RP_BK){
        const int i = (int)(q / plane);
        const size_t r = q - (size_t)i * plane;
        const int j = (int)(r / pitch), k = (int)(r - (size_t)j * pitch);
        if(k >= km) continue;
        double v = (double)f[q];
        if(lo) v += (double)lo[q];
        const double d = fabs(v - fn[q]);
        if(!(d >= 0.0)) continue;
        const Cand c{ d, ((long long)i * jm + j) * km + k };
        if(better_last(c, best)) best = c;
    }
    block_reduce<1>(best, unused);
    if(threadIdx.x == 0) parts[blockIdx.x] = Part{ best, unused };
}

// a field's value (hi + lo in the refined form) as the n-level copy of the report
template<typename T>
__global__ __launch_bounds__(256) void k_snap(const T* __restrict__ f, const float* __restrict__ lo,
        double* __restrict__ out, size_t n){
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if(q >= n) return;
    double v = (double)f[q];
    if(lo) v += (double)lo[q];
    out[q] = v;
}

template<typename T>
__global__ void k_first(const T* __restrict__ f, const float* __restrict__ lo, double* __restrict__ out){
    double v = (double)f[0];
    if(lo) v += (double)lo[0];
    *out = v;
}

template<typename T>
__global__ __launch_bounds__(RP_BK) void k_slice(const T* __restrict__ f, const float* __restrict__ lo, int axis, int index,
        int im, int jm, int km, int pitch, double* __restrict__ dst){
    const int e = blockIdx.x * RP_BK + threadIdx.x;
    int i, j, k;
    if(axis == 0){ if(e >= jm * km) return; i = index; j = e / km; k = e - j * km; }
    else if(axis == 1){ if(e >= im * km) return; j = index; i = e / km; k = e - i * km; }
    else { if(e >= im * jm) return; k = index; i = e / jm; j = e - i * jm; }
    const size_t q = ((size_t)i * jm + j) * pitch + k;
    double v = (double)f[q];
    if(lo) v += (double)lo[q];
    dst[e] = v;
}

// report buffer: [RP_BLOCKS] partials, one result, one double, then the slice
Part* parts_of(atjup_ctx* c){ return (Part*)c->report.p; }
Part* result_of(atjup_ctx* c){ return parts_of(c) + RP_BLOCKS; }
double* first_of(atjup_ctx* c){ return (double*)(result_of(c) + 1); }
double* slice_of(atjup_ctx* c){ return first_of(c) + 2; }

template<typename T>
int do_minmax3d(atjup_ctx* c, int field){
    const T* f = (const T*)c->cur[field].p;
    const float* lo = field < ATJUP_N_PROG ? (const float*)c->lo[field].p : nullptr;
    hipLaunchKernelGGL(k_minmax3d<T>, dim3(RP_BLOCKS), dim3(RP_BK), 0, c->stream, f, lo, c->im, c->jm, c->km, c->pitch, parts_of(c));
    hipLaunchKernelGGL(k_first<T>, dim3(1), dim3(1), 0, c->stream, f, lo, first_of(c));
    hipLaunchKernelGGL(k_reduce_parts<0>, dim3(1), dim3(RP_BK), 0, c->stream, (const Part*)parts_of(c), RP_BLOCKS,
        (const double*)first_of(c), result_of(c));
    return check_launch();
}

template<typename T>
int do_continuity(atjup_ctx* c){
    const float* lu = (const float*)c->lo[ATJUP_F_U].p;
    const float* lv = (const float*)c->lo[ATJUP_F_V].p;
    const float* lw = (const float*)c->lo[ATJUP_F_W].p;
    hipLaunchKernelGGL(k_continuity<T>, dim3(RP_BLOCKS), dim3(RP_BK), 0, c->stream,
        (const T*)c->cur[ATJUP_F_U].p, (const T*)c->cur[ATJUP_F_V].p, (const T*)c->cur[ATJUP_F_W].p, lu, lv, lw,
        (const double*)c->tabs.p, c->im, c->jm, c->km, c->pitch, c->g.dr, c->g.dthe, c->g.dphi, parts_of(c));
    hipLaunchKernelGGL(k_reduce_parts<1>, dim3(1), dim3(RP_BK), 0, c->stream, (const Part*)parts_of(c), RP_BLOCKS,
        (const double*)nullptr, result_of(c));
    return check_launch();
}

template<typename T>
int do_slice(atjup_ctx* c, int field, int axis, int index, int n){
    const float* lo = field < ATJUP_N_PROG ? (const float*)c->lo[field].p : nullptr;
    hipLaunchKernelGGL(k_slice<T>, dim3(nblk(n, RP_BK)), dim3(RP_BK), 0, c->stream, (const T*)c->cur[field].p, lo,
        axis, index, c->im, c->jm, c->km, c->pitch, slice_of(c));
    return check_launch();
}

// the report's field order (Reporting.h:356-369)
const int kSteadyFields[ATJUP_N_STEADY] = { ATJUP_F_P_DYN, ATJUP_F_U, ATJUP_F_V, ATJUP_F_W, ATJUP_F_T, ATJUP_F_H2O,
    ATJUP_F_H2O_CLOUD, ATJUP_F_H2O_ICE, ATJUP_F_NH3, ATJUP_F_NH3_CLOUD, ATJUP_F_NH3_ICE, ATJUP_F_NH4SH };
double* steady_copy(atjup_ctx* c, int e){ return (double*)c->steady.p + (size_t)e * c->n3; }
Part* steady_results(atjup_ctx* c){ return (Part*)(steady_copy(c, ATJUP_N_STEADY)); }

template<typename T>
int do_steady_mark(atjup_ctx* c, unsigned entries){
    for(int e = 0; e < ATJUP_N_STEADY; e++){
        if(!(entries & (1u << e))) continue;
        const int f = kSteadyFields[e];
        const float* lo = f < ATJUP_N_PROG ? (const float*)c->lo[f].p : nullptr;
        hipLaunchKernelGGL(k_snap<T>, dim3(nblk(c->n3, 256)), dim3(256), 0, c->stream, (const T*)c->cur[f].p, lo,
            steady_copy(c, e), c->n3);
    }
    return check_launch();
}

template<typename T>
int do_steady(atjup_ctx* c){
    for(int e = 0; e < ATJUP_N_STEADY; e++){
        const int f = kSteadyFields[e];
        const float* lo = f < ATJUP_N_PROG ? (const float*)c->lo[f].p : nullptr;
        hipLaunchKernelGGL(k_steady<T>, dim3(RP_BLOCKS), dim3(RP_BK), 0, c->stream, (const T*)c->cur[f].p, lo,
            (const double*)steady_copy(c, e), c->im, c->jm, c->km, c->pitch, parts_of(c));
        hipLaunchKernelGGL(k_reduce_parts<1>, dim3(1), dim3(RP_BK), 0, c->stream, (const Part*)parts_of(c), RP_BLOCKS,
            (const double*)nullptr, steady_results(c) + e);
    }
    return check_launch();
}

#define RDISPATCH(fn, ...) ((c->precision == ATJUP_PRECISION_FP32_REF) ? fn<float>(__VA_ARGS__) : fn<double>(__VA_ARGS__))

int fetch_result(atjup_ctx* c, Part& r){
    ATJUP_HIP_CHECK(hipMemcpyAsync(&r, result_of(c), sizeof(Part), hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // namespace

size_t report_bytes(int im, int jm, int km){
    const size_t slice = std::max((size_t)jm * km, std::max((size_t)im * km, (size_t)im * jm));
    return (RP_BLOCKS + 1) * sizeof(Part) + 2 * sizeof(double) + slice * sizeof(double);
}

int steady_mark(atjup_ctx* c, unsigned entries){
    if(!c->steady_on || !c->steady.p) return 0;
    return RDISPATCH(do_steady_mark, c, entries);
}

// ---- the allocation-free search (atjup_hip_search_minmax_3d_noalloc) ---------------------------
// Its partial results live in this file-scope device array instead of a context's report buffer:
// RP_BLOCKS partials, the combined result and the starting element. Each GPU holds its own copy
// for the life of the process; host code reaches it only through the address hipGetSymbolAddress
// returns on that device (the array's name is a host-side stand-in, not a device address).
constexpr size_t NOALLOC_SCRATCH_BYTES = (RP_BLOCKS + 1) * sizeof(Part) + 2 * sizeof(double);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on the current device, looked up once per device
static int noalloc_scratch(void** out){
    int dev = -1;
    ATJUP_HIP_CHECK(hipGetDevice(&dev));
    if(dev < 0 || dev >= NOALLOC_MAX_DEVICES){ set_error("device index above 63"); return ATJUP_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if(!g_scratch_ptr[dev]){
        void* p = nullptr;
        ATJUP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if(!p){ set_error("hipGetSymbolAddress returned no address for the scratch array"); return ATJUP_ERR_HIP; }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace atjup

using namespace atjup;

extern "C" {

int atjup_hip_search_minmax_3d(atjup_ctx* c, int field, double* out, int* loc){
    if(!c || !out || !loc){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(field < 0 || field >= ATJUP_F_END || !c->cur[field].p){ set_error("field id"); return ATJUP_ERR_ARG; }
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("the tangent form has no reporting"); return ATJUP_ERR_PROFILE; }
    if(int rc = RDISPATCH(do_minmax3d, c, field)) return rc;
    Part r;
    if(int rc = fetch_result(c, r)) return rc;
    const long long o[2] = { r.mx.o, r.mn.o };
    out[0] = r.mx.v; out[1] = r.mn.v;
    for(int s = 0; s < 2; s++){                              // position (j * km + k) * im + i
        loc[3 * s + 0] = (int)(o[s] % c->im);
        loc[3 * s + 2] = (int)((o[s] / c->im) % c->km);
        loc[3 * s + 1] = (int)(o[s] / ((long long)c->im * c->km));
    }
    return 0;
}

// searchMinMax_3D over a caller's device field with no context and no allocation: the partial
// results go to the file-scope scratch array. Calls that overlap on one GPU share that array and
// race; each call writes every partial it reads, so nothing carries over between calls.
int atjup_hip_search_minmax_3d_noalloc(const double* field, int im, int jm, int km, int pitch,
                                       void* stream, double* out, int* loc){
    if(!field || !out || !loc){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(im < 1 || jm < 1 || km < 1 || pitch < km){ set_error("grid: im, jm, km >= 1 and pitch >= km"); return ATJUP_ERR_ARG; }
    constexpr size_t MAX_CELLS = (size_t)1 << 36;
    const size_t cells = (size_t)im * jm * pitch;
    if(im > 4096 || jm > 65536 || pitch > 65536 || cells > MAX_CELLS){
        set_error("search_minmax_3d_noalloc: the field exceeds the maximum (im <= 4096, jm <= 65536, "
                  "pitch <= 65536, im * jm * pitch <= 2^36)");
        return ATJUP_ERR_ARG;
    }
    void* s = nullptr;
    if(int rc = noalloc_scratch(&s)) return rc;
    Part* parts = (Part*)s;
    Part* res = parts + RP_BLOCKS;
    double* first = (double*)(res + 1);
    hipStream_t st = (hipStream_t)stream;
    const int nb = (int)std::min<size_t>(RP_BLOCKS, (cells + RP_BK - 1) / RP_BK);
    hipLaunchKernelGGL(k_minmax3d<double>, dim3(nb), dim3(RP_BK), 0, st, field, (const float*)nullptr,
        im, jm, km, pitch, parts);
    hipLaunchKernelGGL(k_first<double>, dim3(1), dim3(1), 0, st, field, (const float*)nullptr, first);
    hipLaunchKernelGGL(k_reduce_parts<0>, dim3(1), dim3(RP_BK), 0, st, (const Part*)parts, nb,
        (const double*)first, res);
    if(int rc = check_launch()) return rc;
    Part r;
    ATJUP_HIP_CHECK(hipMemcpyAsync(&r, res, sizeof(Part), hipMemcpyDeviceToHost, st));
    ATJUP_HIP_CHECK(hipStreamSynchronize(st));
    const long long o[2] = { r.mx.o, r.mn.o };
    out[0] = r.mx.v; out[1] = r.mn.v;
    for(int q = 0; q < 2; q++){                              // position (j * km + k) * im + i
        loc[3 * q + 0] = (int)(o[q] % im);
        loc[3 * q + 2] = (int)((o[q] / im) % km);
        loc[3 * q + 1] = (int)(o[q] / ((long long)im * km));
    }
    return 0;
}

int atjup_hip_search_minmax_2d(atjup_ctx* c, int map, double* out, int* loc){
    if(!c || !out || !loc){ set_error("null argument"); return ATJUP_ERR_ARG; }
    const double* f;
    if(map >= 0 && map < 5) f = (const double*)c->precip_srf.p + (size_t)map * c->jm * c->km;
    else if(map == 5) f = (const double*)c->vel_star.p;
    else { set_error("map id: 0-4 the surface precipitation maps, 5 vel_star"); return ATJUP_ERR_ARG; }
    hipLaunchKernelGGL(k_minmax2d, dim3(256), dim3(RP_BK), 0, c->stream, f, c->jm, c->km, parts_of(c));
    hipLaunchKernelGGL(k_reduce_parts<0>, dim3(1), dim3(RP_BK), 0, c->stream, (const Part*)parts_of(c), 256, f, result_of(c));
    if(int rc = check_launch()) return rc;
    Part r;
    if(int rc = fetch_result(c, r)) return rc;
    out[0] = r.mx.v; out[1] = r.mn.v;
    const long long o[2] = { r.mx.o, r.mn.o };
    for(int s = 0; s < 2; s++){                              // position j * km + k, the first element -1
        const long long p = o[s] < 0 ? 0 : o[s];
        loc[2 * s + 0] = (int)(p / c->km);
        loc[2 * s + 1] = (int)(p % c->km);
    }
    return 0;
}

int atjup_hip_steady_track(atjup_ctx* c, int on){
    if(!c){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("the tangent form has no reporting"); return ATJUP_ERR_PROFILE; }
    if(on && !c->steady.p){
        const size_t bytes = (size_t)ATJUP_N_STEADY * c->n3 * sizeof(double) + ATJUP_N_STEADY * sizeof(Part);
        if(hipMalloc(&c->steady.p, bytes) != hipSuccess){ c->steady.p = nullptr; set_error("hipMalloc failed (steadyQuery tracking)"); return ATJUP_ERR_NOMEM; }
        c->steady.bytes = bytes;
    }
    c->steady_on = on != 0;
    // the n level as it stands: the copies a query before the next iteration differences against
    return on ? steady_mark(c, (1u << ATJUP_N_STEADY) - 1) : 0;
}

int atjup_hip_steady_changes(atjup_ctx* c, double* values, int* cells){
    if(!c || !values || !cells){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(!c->steady_on || !c->steady.p){ set_error("steadyQuery tracking is off (atjup_hip_steady_track)"); return ATJUP_ERR_STATE; }
    if(int rc = RDISPATCH(do_steady, c)) return rc;
    Part r[ATJUP_N_STEADY];
    ATJUP_HIP_CHECK(hipMemcpyAsync(r, steady_results(c), sizeof(r), hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    for(int e = 0; e < ATJUP_N_STEADY; e++){
        const long long o = r[e].mx.o < 0 ? 0 : r[e].mx.o;   // position (i * jm + j) * km + k
        values[e] = r[e].mx.o < 0 ? 0.0 : r[e].mx.v;
        cells[3 * e + 2] = (int)(o % c->km);
        cells[3 * e + 1] = (int)((o / c->km) % c->jm);
        cells[3 * e + 0] = (int)(o / ((long long)c->km * c->jm));
    }
    return 0;
}

int atjup_hip_continuity_residual(atjup_ctx* c, double* value, int* loc){
    if(!c || !value || !loc){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("the tangent form has no reporting"); return ATJUP_ERR_PROFILE; }
    if(int rc = RDISPATCH(do_continuity, c)) return rc;
    Part r;
    if(int rc = fetch_result(c, r)) return rc;
    if(r.mx.o < 0){ *value = 0.0; loc[0] = loc[1] = loc[2] = -1; return 0; }
    *value = r.mx.v;
    loc[2] = (int)(r.mx.o % c->km);
    loc[1] = (int)((r.mx.o / c->km) % c->jm);
    loc[0] = (int)(r.mx.o / ((long long)c->km * c->jm));
    return 0;
}

int atjup_hip_download_slice(atjup_ctx* c, int field, int axis, int index, double* dst){
    if(!c || !dst){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(field < 0 || field >= ATJUP_F_END || !c->cur[field].p){ set_error("field id"); return ATJUP_ERR_ARG; }
    if(c->precision == ATJUP_PRECISION_TANGENT){ set_error("the tangent form has no reporting"); return ATJUP_ERR_PROFILE; }
    const int ext[3] = { c->im, c->jm, c->km };
    if(axis < 0 || axis > 2 || index < 0 || index >= ext[axis]){ set_error("slice axis or index"); return ATJUP_ERR_ARG; }
    const int n = axis == 0 ? c->jm * c->km : axis == 1 ? c->im * c->km : c->im * c->jm;
    if(int rc = RDISPATCH(do_slice, c, field, axis, index, n)) return rc;
    ATJUP_HIP_CHECK(hipMemcpyAsync(dst, slice_of(c), (size_t)n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    ATJUP_HIP_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"

// The device math library's FP64 exp, log, pow and tanh, the transcendental functions the time step's
// kernels call, evaluated elementwise so that their distance from the host's libm can be measured.
__global__ void k_libm_probe(int fn, const double* __restrict__ x, const double* __restrict__ y,
                             double* __restrict__ out, long n){
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if(q >= n) return;
    const double a = x[q];
    out[q] = fn == 0 ? exp(a) : fn == 1 ? log(a) : fn == 2 ? pow(a, y[q]) : tanh(a);
}

extern "C" int atjup_hip_libm_probe(int fn, const double* x, const double* y, double* out, long n){
    if(!x || !out || (fn == 2 && !y)){ set_error("null argument"); return ATJUP_ERR_ARG; }
    if(fn < 0 || fn > 3 || n < 1 || n > (1L << 26)){ set_error("fn 0-3 (exp, log, pow, tanh), 1 <= n <= 2^26"); return ATJUP_ERR_ARG; }
    double* d = nullptr;
    const size_t b = (size_t)n * sizeof(double);
    ATJUP_HIP_CHECK(hipMalloc(&d, 3 * b));
    int rc = 0;
    if(hipMemcpy(d, x, b, hipMemcpyHostToDevice) != hipSuccess ||
       (fn == 2 && hipMemcpy(d + n, y, b, hipMemcpyHostToDevice) != hipSuccess)) rc = ATJUP_ERR_HIP;
    if(!rc){
        hipLaunchKernelGGL(k_libm_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, d, d + n, d + 2 * n, n);
        if(hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, b, hipMemcpyDeviceToHost) != hipSuccess)
            rc = ATJUP_ERR_HIP;
    }
    if(rc) set_error("libm_probe: HIP call failed");
    hipFree(d);
    return rc;
}
