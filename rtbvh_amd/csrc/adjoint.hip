// Adjoint of the ray/triangle hit (no counterpart in the reference, whose
// Moller-Trumbore test, src/builders/spatial_sah.rs:132-163, is forward only).
//
// With the hit triangle fixed by the forward pass, (t, u, v) solves
//     o + t d = (1 - u - v) v0 + u v1 + v v2,   i.e.   M x = s
// with M = [-d | e1 | e2], x = (t, u, v), s = o - v0, e1 = v1 - v0,
// e2 = v2 - v0.  For an upstream gradient g = dL/dx the adjoint is
//     lambda = M^-T g,   dL/ds = lambda,   dL/dM = -lambda x^T,
// which gives
//     dL/do  = lambda                 dL/dd  = t lambda
//     dL/dv0 = -(1 - u - v) lambda    dL/dv1 = -u lambda    dL/dv2 = -v lambda.
// The rows of M^-1 are f (e1 x e2), f (d x e2), f (e1 x d) with f = 1 / a,
// a = e1 . (d x e2) = det M (the same a as the forward test), so
//     lambda = f (g_t (e1 x e2) + g_u (d x e2) + g_v (e1 x d)).
// Vertex gradients of triangles hit by several rays are summed with float
// atomics (order-dependent in the last bits); ray gradients are written.
#include "internal.hpp"

namespace rtbvh {

namespace {

__device__ __forceinline__ void cross3(float ax, float ay, float az, float bx, float by, float bz,
                                       float& cx, float& cy, float& cz) {
    cx = ay * bz - by * az;
    cy = az * bx - bz * ax;
    cz = ax * by - bx * ay;
}

__global__ void adjoint_kernel(const float* verts, uint32_t prim_count, const DevRay* rays,
                               const DevHit* hits, uint64_t ray_count, const float* grad_tuv,
                               float* grad_verts, float* grad_origin, float* grad_dir) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= ray_count) return;
    const float4 h = reinterpret_cast<const float4*>(hits)[i];
    const uint32_t prim = __float_as_uint(h.w);
    float lx = 0.f, ly = 0.f, lz = 0.f;
    const float t = h.x, u = h.y, v = h.z;
    if (prim < prim_count) {
        const float4* r4 = reinterpret_cast<const float4*>(rays + i);
        const float4 rb = r4[1];
        const float dx = rb.x, dy = rb.y, dz = rb.z;
        const float* p = verts + size_t(prim) * 9;
        const float e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
        const float e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
        float hx, hy, hz, nx, ny, nz, kx, ky, kz;
        cross3(dx, dy, dz, e2x, e2y, e2z, hx, hy, hz);     // d x e2
        cross3(e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz);  // e1 x e2
        cross3(e1x, e1y, e1z, dx, dy, dz, kx, ky, kz);     // e1 x d
        const float a = (e1x * hx + e1y * hy) + e1z * hz;
        const float f = 1.0f / a;
        const float gt = grad_tuv[i * 3], gu = grad_tuv[i * 3 + 1], gv = grad_tuv[i * 3 + 2];
        lx = f * (gt * nx + gu * hx + gv * kx);
        ly = f * (gt * ny + gu * hy + gv * ky);
        lz = f * (gt * nz + gu * hz + gv * kz);
        if (grad_verts) {
            float* gp = grad_verts + size_t(prim) * 9;
            const float w0 = -(1.0f - u - v), w1 = -u, w2 = -v;
            atomicAdd(gp + 0, w0 * lx);
            atomicAdd(gp + 1, w0 * ly);
            atomicAdd(gp + 2, w0 * lz);
            atomicAdd(gp + 3, w1 * lx);
            atomicAdd(gp + 4, w1 * ly);
            atomicAdd(gp + 5, w1 * lz);
            atomicAdd(gp + 6, w2 * lx);
            atomicAdd(gp + 7, w2 * ly);
            atomicAdd(gp + 8, w2 * lz);
        }
    }
    if (grad_origin) {
        grad_origin[i * 3] = lx;
        grad_origin[i * 3 + 1] = ly;
        grad_origin[i * 3 + 2] = lz;
    }
    if (grad_dir) {
        const bool hit = prim < prim_count;
        grad_dir[i * 3] = hit ? t * lx : 0.f;
        grad_dir[i * 3 + 1] = hit ? t * ly : 0.f;
        grad_dir[i * 3 + 2] = hit ? t * lz : 0.f;
    }
}

}  // namespace

hipError_t launch_adjoint(const float* verts, uint32_t prim_count, const DevRay* rays,
                          const DevHit* hits, uint64_t ray_count, const float* grad_tuv,
                          float* grad_verts, float* grad_origin, float* grad_dir,
                          hipStream_t stream) {
    if (ray_count == 0) return hipSuccess;
    const uint64_t blocks = (ray_count + 255) / 256;
    hipLaunchKernelGGL(adjoint_kernel, dim3(uint32_t(blocks)), dim3(256), 0, stream, verts,
                       prim_count, rays, hits, ray_count, grad_tuv, grad_verts, grad_origin,
                       grad_dir);
    return hipGetLastError();
}

}  // namespace rtbvh
