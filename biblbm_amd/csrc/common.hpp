// Shared definitions for the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>
#include <string>

#include "../../include/biblbm.h"

namespace bl {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define BL_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::bl::fail(BL_ERR_HIP, std::string(#call) + ": " +             \
                                              hipGetErrorString(e_));             \
    } while (0)

#define BL_CHECK_LAUNCH() BL_HIP(hipGetLastError())

// ------------------------------------------------------- D3Q19 (Palabos)
// Velocity set and weights of the reference descriptor
// (src/latticeBoltzmann/nearestNeighborLattices3D.hh:165-199): i and i+9 are
// opposite; i = 1..9 point into the negative half space.
constexpr int Q = 19;
__host__ __device__ constexpr int cx(int i)
{
    constexpr int v[19] = {0, -1, 0, 0, -1, -1, -1, -1, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
    return v[i];
}
__host__ __device__ constexpr int cy(int i)
{
    constexpr int v[19] = {0, 0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 1, 0, 1, -1, 0, 0, 1, 1};
    return v[i];
}
__host__ __device__ constexpr int cz(int i)
{
    constexpr int v[19] = {0, 0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 1, 0, 0, 1, -1, 1, -1};
    return v[i];
}
__host__ __device__ constexpr int ci(int i, int d) { return d == 0 ? cx(i) : (d == 1 ? cy(i) : cz(i)); }
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : (i <= 9 ? i + 9 : i - 9); }
__host__ __device__ constexpr double tw(int i)
{
    return i == 0 ? 1.0 / 3.0 : ((i <= 3 || (i >= 10 && i <= 12)) ? 1.0 / 18.0 : 1.0 / 36.0);
}

// FP16 "shifted" storage: h_i = half(fBar_i / t_i * 2^10).
constexpr double kF16Scale = 1024.0;

// Tiles of 4x4x4 cells carry the IBM force field.
constexpr int kTile = 4;

// Mask codes.
enum : unsigned char { M_FLUID = 0, M_BB = 1, M_NONE = 2, M_VWALL = 3 };

}  // namespace bl
