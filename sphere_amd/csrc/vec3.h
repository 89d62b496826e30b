// vec3.h -- 3-vector arithmetic for the device code of the DEM step.
//
// The operators keep the operand order of the reference's vector_arithmetic.h
// (src/vector_arithmetic.h:36-417): every component is one IEEE operation in
// the written order, so the FP64 kernels reproduce the CPU oracle bit for bit
// (built without FP contraction).
#ifndef SPHERE_VEC3_H_
#define SPHERE_VEC3_H_

#include <hip/hip_runtime.h>

namespace sph {

template <class T> struct V3 { T x, y, z; };

template <class T> __device__ __forceinline__ V3<T> mk(T x, T y, T z) { V3<T> r; r.x = x; r.y = y; r.z = z; return r; }
template <class T> __device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ __forceinline__ T vlen(V3<T> a) { return sqrt(dot(a, a)); }
template <class T> __device__ __forceinline__ V3<T> cross(V3<T> a, V3<T> b) {
    return mk<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> __device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return mk<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return mk<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
// scalar * vector and vector * scalar keep the operand order of
// src/vector_arithmetic.h:150-157
template <class T> __device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return mk<T>(s * a.x, s * a.y, s * a.z); }
template <class T> __device__ __forceinline__ V3<T> operator*(V3<T> a, T s) { return mk<T>(a.x * s, a.y * s, a.z * s); }
template <class T> __device__ __forceinline__ V3<T> operator/(V3<T> a, T s) { return mk<T>(a.x / s, a.y / s, a.z / s); }
template <class T> __device__ __forceinline__ V3<T> operator-(V3<T> a) { return mk<T>(-a.x, -a.y, -a.z); }

__device__ __forceinline__ V3<double> xyz(double4 a) { return mk<double>(a.x, a.y, a.z); }
template <class T> __device__ __forceinline__ V3<T> cvt(V3<double> a) { return mk<T>((T)a.x, (T)a.y, (T)a.z); }

}  // namespace sph

#endif
