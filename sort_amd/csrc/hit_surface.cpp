// Hit reconstruction on the host: the SurfaceInteraction fields SORT fills for a closest hit,
// from the (t, prim, u, v) record the traversal returns and the accelerator's own copy of the
// geometry. Triangles follow setupIntersection (simd/simd_triangle.h:311-341), quads the hit
// branch of Quad::GetIntersect (shape/quad.cpp:66-96), each float operation in the reference's
// order (the library is compiled without contraction), so the fields are SORT's floats.
#include <cmath>
#include <cstring>

#include "../../include/sortgpu.h"
#include "api_internal.h"

namespace {

struct F3 {
    float x, y, z;
};
F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
F3 mul(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
// cross (math/vector3.h:151-155)
F3 cross(F3 a, F3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float length(F3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
// normalize(v): v unchanged when its length is 0, else v / len (math/vector3.h:163-168)
F3 normalize(F3 a) {
    const float len = length(a);
    if (len == 0.0f) return a;
    return {a.x / len, a.y / len, a.z / len};
}
// Vector3::Normalize(): the same rule in place (math/vector3.h:104-109)
F3 normalized(F3 a) {
    const float len = length(a);
    if (len != 0) a = {a.x / len, a.y / len, a.z / len};
    return a;
}
F3 load(const float* a, uint32_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
void store(float* o, F3 a) {
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
}
// Matrix::TransformVector of a row-major 4x4 (math/matrix.cpp:85-92)
F3 xform_vector(const float* m, F3 v) {
    return {v.x * m[0] + v.y * m[1] + v.z * m[2], v.x * m[4] + v.y * m[5] + v.z * m[6],
            v.x * m[8] + v.y * m[9] + v.z * m[10]};
}

}  // namespace

extern "C" int sortgpu_hit_surface(const sortgpu_accel* acc, const sortgpu_ray* ray, const sortgpu_hit* hit,
                                   const float* normals, const float* tangents, const float* texcoords,
                                   sortgpu_surface* out) {
    using sortgpu::set_error;
    if (!acc || !ray || !hit || !out) {
        set_error(acc ? acc->ctx : nullptr, "sortgpu_hit_surface: null argument");
        return SORTGPU_ERR_ARG;
    }
    const uint32_t ntris = (uint32_t)(acc->idx.size() / 3);
    const uint32_t nquads = (uint32_t)acc->quads.size();
    if (hit->prim == 0xFFFFFFFFu || hit->prim >= ntris + nquads) {
        set_error(acc->ctx, "sortgpu_hit_surface: primitive %u is not a hit of this accelerator", hit->prim);
        return SORTGPU_ERR_ARG;
    }
    std::memset(out, 0, sizeof(*out));
    const F3 o = {ray->ox, ray->oy, ray->oz}, d = {ray->dx, ray->dy, ray->dz};
    const float t = hit->t;
    store(out->intersect, add(o, mul(t, d)));  // Ray::operator()(t) = o + t * d (math/ray.h:55-57)
    out->t = t;
    store(out->view, neg(d));
    out->prim = hit->prim;
    if (hit->prim < ntris) {
        const uint32_t* ix = &acc->idx[3 * (size_t)hit->prim];
        const float u = hit->u, v = hit->v;
        const float w = 1 - u - v;
        const float* pos = acc->pos.data();
        const F3 p0 = load(pos, ix[0]), p1 = load(pos, ix[1]), p2 = load(pos, ix[2]);
        store(out->gnormal, normalize(cross(sub(p2, p0), sub(p1, p0))));
        if (normals)
            store(out->normal,
                  normalized(add(add(mul(w, load(normals, ix[0])), mul(u, load(normals, ix[1]))),
                                 mul(v, load(normals, ix[2])))));
        if (tangents)
            store(out->tangent,
                  normalized(add(add(mul(w, load(tangents, ix[0])), mul(u, load(tangents, ix[1]))),
                                 mul(v, load(tangents, ix[2])))));
        if (texcoords) {
            const float* a = texcoords + 2 * (size_t)ix[0];
            const float* b = texcoords + 2 * (size_t)ix[1];
            const float* c = texcoords + 2 * (size_t)ix[2];
            out->u = (w * a[0] + u * b[0]) + v * c[0];
            out->v = (w * a[1] + u * b[1]) + v * c[1];
        }
        out->is_quad = 0;
    } else {
        const sortgpu_quad& q = acc->quads[hit->prim - ntris];
        float inv_t[16];  // TransformNormal = invMatrix.Transpose().TransformVector (math/transform.h:54-56)
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) inv_t[4 * r + c] = q.inv[4 * c + r];
        const F3 n = xform_vector(inv_t, F3{0.0f, 1.0f, 0.0f});
        store(out->normal, n);
        store(out->gnormal, n);
        store(out->tangent, xform_vector(q.mat, F3{0.0f, 0.0f, 1.0f}));
        out->is_quad = 1;
    }
    return SORTGPU_OK;
}
