// Device helpers shared by the kernel translation units of libatjup_hip.so.
#pragma once
#include "atjup_internal.h"

#include <type_traits>

// 1: the sweep sequence of a pressure solve is captured once into a HIP graph and replayed
#ifndef ATJUP_SWEEP_GRAPH
#define ATJUP_SWEEP_GRAPH 1
#endif

namespace atjup {

// The FP32-refined form carries t, u, v, w and the species as FP32 pairs hi + lo; lo is null for
// the other forms and fields. A value written through pair_put is split so that hi + lo holds it
// to about 2^-48 of its size.
// In the FP32 storage the low word is loaded without a branch (from the value's own array where
// the field has none, and then not added): a load under a branch makes the wavefront wait at the
// branch's end for every load it has in flight, which serialises a kernel's loads field by field.
template<typename T>
__device__ __forceinline__ double pair_get(const T* hi, const float* lo, size_t q){
    if constexpr(std::is_same<T, float>::value){
        const float h = hi[q];
        const float l = (lo ? lo : hi)[q];
        return (double)h + (lo ? (double)l : 0.0);
    } else {
        return lo ? (double)hi[q] + (double)lo[q] : (double)hi[q];
    }
}
template<typename T>
__device__ __forceinline__ void pair_put(T* hi, float* lo, size_t q, double v){
    if(lo){
        const float h = (float)v;
        hi[q] = (T)h;
        lo[q] = (float)(v - (double)h);
    } else {
        hi[q] = (T)v;
    }
}

__device__ __forceinline__ int xcd_remap(int b, int nb){
    // The dispatcher hands block b to XCD b % 8. Give XCD x the contiguous range of logical
    // blocks [start_x, start_x + count_x) instead, so neighbouring tiles share an L2.
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, n = b >> 3;
    const int start = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return start + n;
}

// the two-element vector of a value type, for loads of two neighbouring cells at once
template<typename T> struct Pair;
template<> struct Pair<double> { using V = double2; };
template<> struct Pair<float>  { using V = float2; };

// An index every active lane of a wavefront holds (a per-(i, j) table entry where a wavefront
// spans one k row), made explicit so that the entry is read once per wavefront through scalar
// loads instead of once per lane through vector loads.
__device__ __forceinline__ int wave_uniform(int v){ return __builtin_amdgcn_readfirstlane(v); }

template<typename T>
__device__ __forceinline__ T safe_clamp_d(T v, T lo, T hi){
    if(!isfinite(v)) return lo;                   // exponent-bit test of RungeKutta_Jup_Turb.cpp:285-290
    return (v < lo) ? lo : ((v > hi) ? hi : v);
}

// bitwise equality (a NaN equals itself, +0 and -0 differ): a store of `a` over `b` changes nothing
__device__ __forceinline__ bool same_bits(double a, double b){
    return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ __forceinline__ bool same_bits(float a, float b){
    return __float_as_int(a) == __float_as_int(b);
}

// std::max / std::min with their exact operand semantics (NaN and signed-zero behaviour included)
template<typename T>
__device__ __forceinline__ T dmax(T a, T b){ return (a < b) ? b : a; }
template<typename T>
__device__ __forceinline__ T dmin(T a, T b){ return (b < a) ? b : a; }

inline int nblk(size_t n, int b){ return (int)((n + b - 1) / b); }

// ---------------------------------------------------------------------------------------------
// Forward-mode derivative: a value with its directional derivative. The dynamics kernels are
// templates on the storage type; instantiated on Dual they compute the tangent-linear step of the
// FP64 path alongside the step itself (comparisons and branches follow the value, so the
// derivative is the one of the branch the forward takes). The adjoint is checked against it.
// ---------------------------------------------------------------------------------------------
struct Dual {
    double v, d;
    Dual() = default;
    __host__ __device__ constexpr Dual(double a) : v(a), d(0.0) {}
    __host__ __device__ constexpr Dual(double a, double b) : v(a), d(b) {}
    __host__ __device__ explicit operator double() const { return v; }
    __host__ __device__ explicit operator float() const { return (float)v; }
};
__host__ __device__ inline Dual operator+(Dual a, Dual b){ return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(Dual a, Dual b){ return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator*(Dual a, Dual b){ return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(Dual a, Dual b){
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ inline Dual operator-(Dual a){ return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator+(Dual a){ return a; }
__host__ __device__ inline Dual operator+(Dual a, double b){ return Dual(a.v + b, a.d); }
__host__ __device__ inline Dual operator+(double a, Dual b){ return Dual(a + b.v, b.d); }
__host__ __device__ inline Dual operator-(Dual a, double b){ return Dual(a.v - b, a.d); }
__host__ __device__ inline Dual operator-(double a, Dual b){ return Dual(a - b.v, -b.d); }
__host__ __device__ inline Dual operator*(Dual a, double b){ return Dual(a.v * b, a.d * b); }
__host__ __device__ inline Dual operator*(double a, Dual b){ return Dual(a * b.v, a * b.d); }
__host__ __device__ inline Dual operator/(Dual a, double b){ return Dual(a.v / b, a.d / b); }
__host__ __device__ inline Dual operator/(double a, Dual b){ const double q = a / b.v; return Dual(q, -q * b.d / b.v); }
__host__ __device__ inline Dual& operator+=(Dual& a, Dual b){ a = a + b; return a; }
__host__ __device__ inline Dual& operator-=(Dual& a, Dual b){ a = a - b; return a; }
__host__ __device__ inline Dual& operator*=(Dual& a, Dual b){ a = a * b; return a; }
#define ATJUP_DUAL_CMP(op) \
    __host__ __device__ inline bool operator op(Dual a, Dual b){ return a.v op b.v; } \
    __host__ __device__ inline bool operator op(Dual a, double b){ return a.v op b; } \
    __host__ __device__ inline bool operator op(double a, Dual b){ return a op b.v; }
ATJUP_DUAL_CMP(<) ATJUP_DUAL_CMP(>) ATJUP_DUAL_CMP(<=) ATJUP_DUAL_CMP(>=) ATJUP_DUAL_CMP(==) ATJUP_DUAL_CMP(!=)
#undef ATJUP_DUAL_CMP
// keep the library's overloads for the built-in types visible beside the Dual ones
using ::isfinite;
using ::sqrt;
using ::fabs;
__device__ inline bool isfinite(Dual a){ return ::isfinite(a.v); }
__device__ inline Dual sqrt(Dual a){ const double r = ::sqrt(a.v); return Dual(r, a.d / (2.0 * r)); }
__device__ inline Dual fabs(Dual a){ return (a.v < 0.0) ? -a : a; }
// exp and log for code templated on double or Dual (the double forms are the library's own)
__device__ __forceinline__ double dual_exp(double a){ return exp(a); }
__device__ __forceinline__ double dual_log(double a){ return log(a); }
__device__ inline Dual dual_exp(Dual a){ const double e = ::exp(a.v); return Dual(e, e * a.d); }
__device__ inline Dual dual_log(Dual a){ return Dual(::log(a.v), a.d / a.v); }

// The type a kernel accumulates and forms its FP64 quantities in: FP64 for the FP64 and FP32
// storage, the derivative-carrying type for the tangent form.
template<typename T> struct acc_type { using type = double; };
template<> struct acc_type<Dual> { using type = Dual; };
template<typename T> using acc_t = typename acc_type<T>::type;
// Metric tables: FP32 for the FP32 storage, FP64 otherwise.
template<typename T> struct geo_type { using type = T; };
template<> struct geo_type<Dual> { using type = double; };
template<typename T> using geo_t = typename geo_type<T>::type;

inline int check_launch(){
    hipError_t e = hipGetLastError();
    if(e != hipSuccess){ set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return ATJUP_ERR_HIP; }
    return 0;
}

}  // namespace atjup
