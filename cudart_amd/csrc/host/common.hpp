// Host-side helpers shared by libprt.so translation units: error state, status
// plumbing and small float3 arithmetic with the operator semantics of
// src/include/linalg.h (component-wise, left to right).
#pragma once

#include <cmath>
#include <cstdint>
#include <string>

#include "prt.h"

namespace prt {

// Thread-local last error text (prt_last_error). Replaces the print-and-exit of
// CUDA_CHECK_RETURN (src/include/cuda_error_check.h:9-21).
int fail(int code, const std::string& msg);
void clear_error();

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 f3(const float* p) { return F3{p[0], p[1], p[2]}; }
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator*(F3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
inline F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float length(F3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline F3 normalize(F3 v) {
    float n = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return {v.x / n, v.y / n, v.z / n};
}

constexpr float kLightsGain = 3.0f;  // LIGHTS_GAIN, src/geometry/scene.cpp:19

}  // namespace prt

#define PRT_TRY_BEGIN try {
#define PRT_TRY_END                                                       \
    }                                                                     \
    catch (const std::bad_alloc&) {                                       \
        return prt::fail(PRT_ERR_OUT_OF_MEMORY, "host allocation failed"); \
    }                                                                     \
    catch (const std::exception& e) {                                     \
        return prt::fail(PRT_ERR_INVALID_ARG, e.what());                  \
    }
