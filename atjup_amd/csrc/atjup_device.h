// Device helpers shared by the kernel translation units of libatjup_hip.so.
#pragma once
#include "atjup_internal.h"

namespace atjup {

__device__ __forceinline__ int xcd_remap(int b, int nb){
    // The dispatcher hands block b to XCD b % 8. Give XCD x the contiguous range of logical
    // blocks [start_x, start_x + count_x) instead, so neighbouring tiles share an L2.
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, n = b >> 3;
    const int start = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
    return start + n;
}

template<typename T>
__device__ __forceinline__ T safe_clamp_d(T v, T lo, T hi){
    if(!isfinite(v)) return lo;                   // exponent-bit test of RungeKutta_Jup_Turb.cpp:285-290
    return (v < lo) ? lo : ((v > hi) ? hi : v);
}

// std::max / std::min with their exact operand semantics (NaN and signed-zero behaviour included)
template<typename T>
__device__ __forceinline__ T dmax(T a, T b){ return (a < b) ? b : a; }
template<typename T>
__device__ __forceinline__ T dmin(T a, T b){ return (b < a) ? b : a; }

inline int nblk(size_t n, int b){ return (int)((n + b - 1) / b); }

inline int check_launch(){
    hipError_t e = hipGetLastError();
    if(e != hipSuccess){ set_error(std::string("kernel launch: ") + hipGetErrorString(e)); return ATJUP_ERR_HIP; }
    return 0;
}

}  // namespace atjup
