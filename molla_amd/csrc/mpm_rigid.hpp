// Rigid-body coupling on the grid (mpm.wgsl rigid_imprint :927-1006,
// mpm_rigid_force_integrate :1174-1226, the rigid mask in grid_update
// :1034-1038 and mpm_implicit_apply :1100-1103).
//
// A body is five float4 rows (RigidDev::body + 5 r):
//   0  center xyz, bounding radius (legacy sphere: its radius)
//   1  linear velocity xyz
//   2  angular velocity xyz
//   3  orientation quaternion xyzw (unit)
//   4  shape: dims xyz, kind w (0 legacy sphere, 1 sphere, 2 box, 3 cylinder
//      with its axle along local y)
// The bodies are either all legacy spheres (set_rigid_spheres: heavy-cell
// stamp, a shell force outside) or all prescribed solids
// (set_prescribed_solids: the sand mass keeps its mass and sticks).
#pragma once

#include "mpm_internal.hpp"

namespace mpm {

struct RigidDev {
    const float4* body;  // 5 rows per body
    double* wrench;      // 8 per body (force xyz, -, torque xyz, -), or null: no reaction this substep
    uint32_t count;
    uint32_t prescribed;
};

constexpr int RIGID_ROWS = 5;
constexpr float RIGID_STAMP_MASS = 1.0e6f;  // legacy spheres: heavy-cell stamp

__device__ __forceinline__ float3 f3(const float4& v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// World position of grid node (ix, iy, iz) (window-relative indices).
__device__ __forceinline__ float3 node_position(const SceneDev& sc, int ix, int iy, int iz) {
    return make_float3(sc.origin[0] + float(ix + sc.window[0]) * sc.dx,
                       sc.origin[1] + float(iy + sc.window[1]) * sc.dx,
                       sc.origin[2] + float(iz + sc.window[2]) * sc.dx);
}

// Whether the node at `arm` from body r's center lies inside it
// (prescribed_solid_contains, mpm.wgsl:927-935, and the legacy distance
// test): the arm is rotated into the body frame by the conjugate quaternion.
__device__ __forceinline__ bool rigid_contains(const float4* b, float3 arm) {
    const float4 shape = b[4];
    if (shape.w == 0.0f) return sqrtf(arm.x * arm.x + arm.y * arm.y + arm.z * arm.z) < b[0].w;
    if (shape.w == 1.0f) return sqrtf(arm.x * arm.x + arm.y * arm.y + arm.z * arm.z) < shape.x;
    const float4 q = b[3];
    const float3 inv = make_float3(-q.x, -q.y, -q.z);
    float3 t = cross3(inv, arm);
    t = make_float3(2.0f * t.x, 2.0f * t.y, 2.0f * t.z);
    const float3 u = cross3(inv, t);
    const float3 local = make_float3(arm.x + q.w * t.x + u.x, arm.y + q.w * t.y + u.y, arm.z + q.w * t.z + u.z);
    if (shape.w == 2.0f) return fabsf(local.x) < shape.x && fabsf(local.y) < shape.y && fabsf(local.z) < shape.z;
    return sqrtf(local.x * local.x + local.z * local.z) < shape.x && fabsf(local.y) < shape.y;
}

// Imprint of one cell (rigid_imprint): the first body containing the node
// gives the cell its velocity.  q = (m, p) of the P2G total, replaced by the
// stamp.  Returns the owning body (-1: none) and, when the cell held sand
// mass, the reaction on that body in force/torque.
__device__ __forceinline__ int rigid_imprint(const RigidDev& rg, const SceneDev& sc, int ix, int iy, int iz,
                                             float dt, float4& q, float3& force, float3& torque) {
    force = torque = make_float3(0.f, 0.f, 0.f);
    const float3 node = node_position(sc, ix, iy, iz);
    for (uint32_t r = 0; r < rg.count; ++r) {
        const float4* b = rg.body + RIGID_ROWS * r;
        const float3 center = f3(b[0]);
        const float3 arm = make_float3(node.x - center.x, node.y - center.y, node.z - center.z);
        if (!rigid_contains(b, arm)) continue;
        const bool solid = b[4].w != 0.0f;
        float3 v = f3(b[1]);
        if (solid) {
            const float3 s = cross3(f3(b[2]), arm);
            v = make_float3(v.x + s.x, v.y + s.y, v.z + s.z);
        }
        const float m_old = q.x;
        if (solid && m_old <= 1.0e-9f) return -1;  // a solid keeps empty cells empty
        const float stamp = solid ? m_old : RIGID_STAMP_MASS;
        const float3 p_old = make_float3(q.y, q.z, q.w);
        q = make_float4(stamp, stamp * v.x, stamp * v.y, stamp * v.z);
        if (m_old > 1.0e-9f) {
            // impulse to the sand m v - p; the body takes -impulse / dt
            float3 imp = make_float3(m_old * v.x - p_old.x, m_old * v.y - p_old.y, m_old * v.z - p_old.z);
            if (solid) imp.y -= m_old * sc.gravity_y * dt;
            const float inv_dt = 1.0f / (solid ? dt : fmaxf(dt, 1.0e-6f));
            force = make_float3(-imp.x * inv_dt, -imp.y * inv_dt, -imp.z * inv_dt);
            torque = cross3(arm, force);
        }
        return int(r);
    }
    return -1;
}

// Whether the node is stamped by a legacy sphere (the implicit sweeps and
// the shell force skip such cells).
__device__ __forceinline__ bool rigid_masked_legacy(const RigidDev& rg, const SceneDev& sc, int ix, int iy, int iz) {
    if (rg.count == 0 || rg.prescribed) return false;
    const float3 node = node_position(sc, ix, iy, iz);
    for (uint32_t r = 0; r < rg.count; ++r) {
        const float4* b = rg.body + RIGID_ROWS * r;
        const float3 arm = make_float3(node.x - b[0].x, node.y - b[0].y, node.z - b[0].z);
        if (rigid_contains(b, arm)) return true;
    }
    return false;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Adds the wave's per-lane reactions to the owners' wrenches: one f64 atomic
// per component per owning body present in the wave.  All 64 lanes call it.
__device__ __forceinline__ void rigid_wrench_add(const RigidDev& rg, int owner, float3 force, float3 torque) {
    uint64_t pending = __ballot(owner >= 0);
    while (pending) {
        const int lead = __ffsll((unsigned long long)pending) - 1;
        const int r = __shfl(owner, lead, 64);
        const bool mine = owner == r;
        const double c[6] = {mine ? double(force.x) : 0.0,  mine ? double(force.y) : 0.0,
                             mine ? double(force.z) : 0.0,  mine ? double(torque.x) : 0.0,
                             mine ? double(torque.y) : 0.0, mine ? double(torque.z) : 0.0};
        double s[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] = wave_sum_f64(c[k]);
        if ((threadIdx.x & 63) == lead) {
            double* w = rg.wrench + 8 * (size_t)r;
            atomicAdd(w + 0, s[0]);
            atomicAdd(w + 1, s[1]);
            atomicAdd(w + 2, s[2]);
            atomicAdd(w + 4, s[3]);
            atomicAdd(w + 5, s[4]);
            atomicAdd(w + 6, s[5]);
        }
        pending &= ~__ballot(mine);
    }
}

}  // namespace mpm
