/* A plain C caller of the drop-in boundary, built by tests/test_abi_c.py with gcc (no HIP
 * compiler, no C++): it includes include/sortgpu.h as a C translation unit, prints the size and
 * every field offset of each public struct (the test compares them with the Python binding's
 * ctypes layouts), and drives the host-side entry points on a host-only context: build a small
 * 8-wide and 4-wide tree, read its info, primitive order and tree dump, and check the status codes
 * of refused calls. Output: one "S name size" line per struct, one "F name.field offset size" line
 * per field, then "R key value..." result lines.
 */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "sortgpu.h"

#define S(T) printf("S %s %zu\n", #T, sizeof(T))
#define F(T, f) printf("F %s.%s %zu %zu\n", #T, #f, offsetof(T, f), sizeof(((T*)0)->f))

static void layouts(void) {
    S(sortgpu_ray);
    F(sortgpu_ray, ox); F(sortgpu_ray, oy); F(sortgpu_ray, oz); F(sortgpu_ray, tmin);
    F(sortgpu_ray, dx); F(sortgpu_ray, dy); F(sortgpu_ray, dz); F(sortgpu_ray, tmax);
    S(sortgpu_hit);
    F(sortgpu_hit, t); F(sortgpu_hit, prim); F(sortgpu_hit, u); F(sortgpu_hit, v);
    S(sortgpu_quad);
    F(sortgpu_quad, mat); F(sortgpu_quad, inv); F(sortgpu_quad, size_x); F(sortgpu_quad, size_y);
    S(sortgpu_build_opts);
    F(sortgpu_build_opts, width); F(sortgpu_build_opts, max_depth); F(sortgpu_build_opts, max_leaf);
    F(sortgpu_build_opts, build_threads);
    S(sortgpu_tuning);
    F(sortgpu_tuning, refill); F(sortgpu_tuning, lds_entries); F(sortgpu_tuning, blocks_per_cu);
    F(sortgpu_tuning, leaf_batch); F(sortgpu_tuning, leaf_min); F(sortgpu_tuning, reorder);
    F(sortgpu_tuning, shade_order); F(sortgpu_tuning, reserved);
    S(sortgpu_material);
    F(sortgpu_material, type); F(sortgpu_material, base); F(sortgpu_material, trans);
    F(sortgpu_material, roughness); F(sortgpu_material, ior); F(sortgpu_material, eta); F(sortgpu_material, k);
    F(sortgpu_material, metallic); F(sortgpu_material, specular); F(sortgpu_material, specular_tint);
    F(sortgpu_material, sheen); F(sortgpu_material, sheen_tint); F(sortgpu_material, clearcoat);
    F(sortgpu_material, clearcoat_gloss);
    S(sortgpu_scene_desc);
    F(sortgpu_scene_desc, normals); F(sortgpu_scene_desc, tangents); F(sortgpu_scene_desc, tri_material);
    F(sortgpu_scene_desc, materials); F(sortgpu_scene_desc, nmaterials); F(sortgpu_scene_desc, quad_radiance);
    F(sortgpu_scene_desc, env_radiance); F(sortgpu_scene_desc, clamp);
    S(sortgpu_camera);
    F(sortgpu_camera, eye); F(sortgpu_camera, target); F(sortgpu_camera, up); F(sortgpu_camera, fov_y);
    S(sortgpu_server_opts);
    F(sortgpu_server_opts, hbm_budget); F(sortgpu_server_opts, wave_paths); F(sortgpu_server_opts, max_active);
    F(sortgpu_server_opts, inflight_waves); F(sortgpu_server_opts, manual); F(sortgpu_server_opts, host_device);
    F(sortgpu_server_opts, fail_alloc_at); F(sortgpu_server_opts, graphs); F(sortgpu_server_opts, reserved);
    S(sortgpu_request);
    F(sortgpu_request, scene); F(sortgpu_request, width); F(sortgpu_request, height); F(sortgpu_request, spp);
    F(sortgpu_request, max_depth); F(sortgpu_request, reserved); F(sortgpu_request, seed);
    F(sortgpu_request, camera); F(sortgpu_request, timeout_s);
    S(sortgpu_request_info);
    F(sortgpu_request_info, state); F(sortgpu_request_info, status); F(sortgpu_request_info, samples_done);
    F(sortgpu_request_info, spp); F(sortgpu_request_info, t_submit); F(sortgpu_request_info, t_admit);
    F(sortgpu_request_info, t_done);
    S(sortgpu_surface);
    F(sortgpu_surface, intersect); F(sortgpu_surface, t); F(sortgpu_surface, gnormal); F(sortgpu_surface, normal);
    F(sortgpu_surface, tangent); F(sortgpu_surface, view); F(sortgpu_surface, u); F(sortgpu_surface, v);
    F(sortgpu_surface, prim); F(sortgpu_surface, is_quad);
    S(sortgpu_scene_file_info);
    F(sortgpu_scene_file_info, threads); F(sortgpu_scene_file_info, spp); F(sortgpu_scene_file_info, image_width);
    F(sortgpu_scene_file_info, image_height); F(sortgpu_scene_file_info, clamping);
    F(sortgpu_scene_file_info, integrator); F(sortgpu_scene_file_info, max_recursive_depth);
    F(sortgpu_scene_file_info, max_bssrdf_bounces); F(sortgpu_scene_file_info, eye); F(sortgpu_scene_file_info, up);
    F(sortgpu_scene_file_info, target); F(sortgpu_scene_file_info, camera_width);
    F(sortgpu_scene_file_info, camera_height); F(sortgpu_scene_file_info, lens_radius);
    F(sortgpu_scene_file_info, sensor_w); F(sortgpu_scene_file_info, sensor_h); F(sortgpu_scene_file_info, aspect_fit);
    F(sortgpu_scene_file_info, pixel_aspect_w); F(sortgpu_scene_file_info, pixel_aspect_h);
    F(sortgpu_scene_file_info, fov); F(sortgpu_scene_file_info, camera_fov_y); F(sortgpu_scene_file_info, has_camera);
    F(sortgpu_scene_file_info, nverts); F(sortgpu_scene_file_info, ntris); F(sortgpu_scene_file_info, nquads);
    F(sortgpu_scene_file_info, nmaterials); F(sortgpu_scene_file_info, pos); F(sortgpu_scene_file_info, normals);
    F(sortgpu_scene_file_info, uv); F(sortgpu_scene_file_info, tangents);
    F(sortgpu_scene_file_info, idx); F(sortgpu_scene_file_info, tri_material);
    F(sortgpu_scene_file_info, quads); F(sortgpu_scene_file_info, quad_radiance);
    F(sortgpu_scene_file_info, materials); F(sortgpu_scene_file_info, env_radiance);
    F(sortgpu_scene_file_info, accelerator); F(sortgpu_scene_file_info, accel_max_depth);
    F(sortgpu_scene_file_info, accel_max_leaf); F(sortgpu_scene_file_info, skipped);
    F(sortgpu_scene_file_info, counts);
}

/* 2 x 2 x 2 grid of unit-spaced small triangles plus one quad light above it */
static int build_and_dump(sortgpu_ctx* ctx, unsigned width) {
    float pos[8 * 3 * 3];
    uint32_t idx[8 * 3];
    int t = 0;
    for (int z = 0; z < 2; z++)
        for (int y = 0; y < 2; y++)
            for (int x = 0; x < 2; x++, t++) {
                const float o[3] = {(float)x, (float)y, (float)z};
                for (int v = 0; v < 3; v++) {
                    pos[9 * t + 3 * v + 0] = o[0] + (v == 1 ? 0.25f : 0.0f);
                    pos[9 * t + 3 * v + 1] = o[1] + (v == 2 ? 0.25f : 0.0f);
                    pos[9 * t + 3 * v + 2] = o[2];
                    idx[3 * t + v] = (uint32_t)(3 * t + v);
                }
            }
    sortgpu_quad q;
    memset(&q, 0, sizeof(q));
    for (int i = 0; i < 4; i++) q.mat[5 * i] = q.inv[5 * i] = 1.0f;
    q.mat[7] = 3.0f;  /* light at y = 3 */
    q.inv[7] = -3.0f;
    q.size_x = q.size_y = 0.5f;
    sortgpu_build_opts o;
    memset(&o, 0, sizeof(o));
    o.width = width;
    o.max_leaf = 2;
    o.build_threads = 1;
    sortgpu_accel* acc = NULL;
    int rc = sortgpu_accel_build(ctx, pos, 24, idx, 8, &q, 1, &o, &acc);
    printf("R build%u %d\n", width, rc);
    if (rc) return rc;
    uint64_t info[8];
    rc = sortgpu_accel_info(acc, info);
    printf("R info%u %d %llu %llu %llu %llu %llu %llu\n", width, rc, (unsigned long long)info[0],
           (unsigned long long)info[1], (unsigned long long)info[2], (unsigned long long)info[3],
           (unsigned long long)info[4], (unsigned long long)info[5]);
    uint32_t order[9];
    rc = sortgpu_accel_prim_order(acc, order, 9);
    printf("R order%u %d", width, rc);
    for (int i = 0; i < 9; i++) printf(" %u", order[i]);
    printf("\n");
    const int64_t n = sortgpu_accel_tree_dump(acc, NULL, 0);
    uint32_t dump[256];
    const int64_t m = n > 0 && n <= 256 ? sortgpu_accel_tree_dump(acc, dump, 256) : -1;
    printf("R dump%u %lld", width, (long long)m);
    for (int64_t i = 0; i < m; i++) printf(" %u", dump[i]);
    printf("\n");
    float box[6];
    rc = sortgpu_accel_bbox(acc, box);
    printf("R bbox%u %d %.9g %.9g %.9g %.9g %.9g %.9g\n", width, rc, box[0], box[1], box[2], box[3], box[4], box[5]);
    /* a host-only context refuses queries and flag checks report nothing */
    sortgpu_ray ray;
    sortgpu_hit hit;
    memset(&ray, 0, sizeof(ray));
    printf("R query%u %d\n", width, sortgpu_trace_closest(acc, &ray, 1, &hit, NULL));
    uint32_t flags = 99;
    rc = sortgpu_accel_errors(acc, NULL, &flags);
    printf("R flags%u %d %u\n", width, rc, flags);
    printf("R valid%u %d\n", width, sortgpu_accel_valid(acc));
    sortgpu_accel_destroy(acc);
    return 0;
}

int main(void) {
    layouts();
    printf("R null_ctx %d\n", sortgpu_ctx_create(0, NULL));
    sortgpu_ctx* ctx = NULL;
    int rc = sortgpu_ctx_create(-1, &ctx);
    printf("R ctx %d\n", rc);
    if (rc) return 1;
    printf("R null_accel_build %d\n", sortgpu_accel_build(ctx, NULL, 0, NULL, 0, NULL, 0, NULL, NULL));
    sortgpu_build_opts bad;
    memset(&bad, 0, sizeof(bad));
    bad.width = 5;
    sortgpu_accel* acc = NULL;
    const float p3[9] = {0, 0, 0, 1, 0, 0, 0, 1, 0};
    const uint32_t i3[3] = {0, 1, 2};
    printf("R bad_width %d\n", sortgpu_accel_build(ctx, p3, 3, i3, 1, NULL, 0, &bad, &acc));
    if (build_and_dump(ctx, 8) || build_and_dump(ctx, 4)) return 1;
    sortgpu_ctx_destroy(ctx);
    return 0;
}
