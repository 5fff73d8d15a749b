// Minimal vector types for the OBJ container and CSCMesh (the subset of the
// reference's include/math/vector.h those headers use: float2, float3,
// float4 with a float3 + w constructor and xyz()).
#pragma once

namespace math {

struct float2 {
  float x = 0.0f, y = 0.0f;
  float2() = default;
  float2(float x_, float y_) : x(x_), y(y_) {}
};

struct float3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  float3() = default;
  float3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};

struct float4 {
  float x = 0.0f, y = 0.0f, z = 0.0f, w = 0.0f;
  float4() = default;
  float4(float x_, float y_, float z_, float w_) : x(x_), y(y_), z(z_), w(w_) {}
  float4(const float3& v, float w_) : x(v.x), y(v.y), z(v.z), w(w_) {}
  float3 xyz() const { return float3(x, y, z); }
};

}  // namespace math
