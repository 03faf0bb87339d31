// rdp_device.hpp -- the relative difference prior's per-pair gradient term (device side).
#pragma once
#include <hip/hip_runtime.h>

namespace omg {

// d/du_j of the RDP pair term with the reference's guards (source/cpp/RDP.h:64-72):
// (u_j - u_k)(gamma |u_j - u_k| + u_j + 3 u_k + eps^2) / ((u_j + u_k + gamma |u_j - u_k|)^2 + eps)
__device__ __forceinline__ float rdp_term(float uj, float uk, float gamma, float epps) {
    const float du = uj - uk;
    const float den = uj + uk + gamma * fabsf(du);
    return du * (gamma * fabsf(du) + uj + 3.f * uk + epps * epps) / (den * den + epps);
}

}  // namespace omg
