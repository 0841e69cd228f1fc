// capi_host.cpp — host-only entry points of the C ABI: BVH build, scene
// loading / generation and camera packing (include/yune_hip.h).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/yune_hip.h"
#include "yn_internal.h"

using namespace yune;

namespace {

int to_c(const HostScene& h, yn_host_scene* out)
{
    std::memset(out, 0, sizeof *out);
    out->n_tri = (int32_t)h.tris.size();
    out->n_mat = (int32_t)h.mats.size();
    out->tris = (yn_triangle*)std::malloc(sizeof(yn_triangle) * std::max<size_t>(1, h.tris.size()));
    out->mats = (yn_material*)std::malloc(sizeof(yn_material) * std::max<size_t>(1, h.mats.size()));
    if (!out->tris || !out->mats) { std::free(out->tris); std::free(out->mats); return YN_ERR_NO_MEMORY; }
    if (!h.tris.empty()) std::memcpy(out->tris, h.tris.data(), sizeof(yn_triangle) * h.tris.size());
    if (!h.mats.empty()) std::memcpy(out->mats, h.mats.data(), sizeof(yn_material) * h.mats.size());
    out->light = h.light;
    out->cam = h.cam;
    return 0;
}
}  // namespace

extern "C" {

int yn_bvh_build(const yn_triangle* tris, int32_t n_tri, int32_t bins, yn_bvh_node** nodes, int32_t* n_nodes)
{
    if (!nodes || !n_nodes || n_tri < 0 || (n_tri > 0 && !tris)) { set_global_error("bad arguments"); return YN_ERR_INVALID; }
    std::vector<yn_bvh_node> out;
    int n = build_sah_bvh(tris, n_tri, bins, 0, out);
    if (n < 0) { set_global_error("BVH build does not terminate on this input"); return YN_ERR_INVALID; }
    *nodes = (yn_bvh_node*)std::malloc(sizeof(yn_bvh_node) * std::max<size_t>(1, out.size()));
    if (!*nodes) return YN_ERR_NO_MEMORY;
    if (!out.empty()) std::memcpy(*nodes, out.data(), sizeof(yn_bvh_node) * out.size());
    *n_nodes = n;
    return 0;
}

int yn_bvh_build_gpu(const yn_triangle* tris, int32_t n_tri, int32_t leaf_max, yn_bvh_node** nodes, int32_t* n_nodes)
{
    if (!nodes || !n_nodes || n_tri < 0 || (n_tri > 0 && !tris) || leaf_max < 1 || leaf_max > 10) {
        set_global_error("bad arguments"); return YN_ERR_INVALID;
    }
    std::vector<yn_bvh_node> out;
    int n = build_gpu_bvh(tris, n_tri, leaf_max, out);
    *nodes = (yn_bvh_node*)std::malloc(sizeof(yn_bvh_node) * std::max<size_t>(1, out.size()));
    if (!*nodes) return YN_ERR_NO_MEMORY;
    if (!out.empty()) std::memcpy(*nodes, out.data(), sizeof(yn_bvh_node) * out.size());
    *n_nodes = n;
    return 0;
}

void yn_free(void* p) { std::free(p); }

int yn_load_obj(const char* path, yn_host_scene* out)
{
    if (!path || !out) { set_global_error("bad arguments"); return YN_ERR_INVALID; }
    HostScene h;
    std::string err;
    int r = load_obj(path, h, err);
    if (r) { set_global_error(err); return r; }
    // reference default camera (Camera.cpp:93-103) and the kernels' constant
    // light quad (udpt.cl:96-104, bdpt.cl same), which every OBJ scene of
    // the reference is rendered with; callers may replace it.
    float side[3] = {1, 0, 0}, up[3] = {0, 1, 0}, look[3] = {0, 0, -1}, eye[3] = {0, 0, 0};
    make_camera(side, up, look, eye, 60.0f, h.cam);
    h.light.pos = {-0.1979f, 0.92f, -3.1972f, 1.f};
    h.light.normal = {0.f, -1.f, 0.f, 0.f};
    h.light.ke = {16.f, 16.f, 16.f, 0.f};
    h.light.edge_l = {0.4f, 0.f, 0.f, 0.f};
    h.light.edge_w = {0.f, 0.f, 0.4f, 0.f};
    return to_c(h, out);
}

int yn_gen_cornell(yn_host_scene* out)
{
    if (!out) return YN_ERR_INVALID;
    HostScene h;
    gen_cornell(h);
    return to_c(h, out);
}

int yn_gen_icosphere_room(uint32_t seed, int32_t n_spheres, int32_t subdiv, yn_host_scene* out)
{
    if (!out || n_spheres < 0 || subdiv < 0 || subdiv > 8) { set_global_error("bad arguments"); return YN_ERR_INVALID; }
    HostScene h;
    gen_icosphere_room(seed, n_spheres, subdiv, h);
    return to_c(h, out);
}

void yn_host_scene_free(yn_host_scene* s)
{
    if (!s) return;
    std::free(s->tris);
    std::free(s->mats);
    s->tris = nullptr;
    s->mats = nullptr;
    s->n_tri = s->n_mat = 0;
}

int yn_camera(const float side[3], const float up[3], const float look_at[3], const float eye[3], float yfov, yn_cam* out)
{
    if (!side || !up || !look_at || !eye || !out) return YN_ERR_INVALID;
    make_camera(side, up, look_at, eye, yfov, *out);
    return 0;
}

// ---- camera controls ----------------------------------------------------

namespace {

void cam_norm3(yn_float4& v)
{
    float d = v.x * v.x + v.y * v.y + v.z * v.z;
    float inv = 1.0f / std::sqrt(d);
    v.x *= inv; v.y *= inv; v.z *= inv;
}

// Rotation by `angle` radians about unit-normalised `axis`, right-handed
// (Rodrigues: R = cos I + sin [k]x + (1 - cos) k k^T), as column vectors.
void rotation(float angle, float ax, float ay, float az, float R[4][4])
{
    float d = 1.0f / std::sqrt(ax * ax + ay * ay + az * az);
    float k[3] = {ax * d, ay * d, az * d};
    float c = std::cos(angle), s = std::sin(angle), t = 1.0f - c;
    float kx[3][3] = {{0, -k[2], k[1]}, {k[2], 0, -k[0]}, {-k[1], k[0], 0}};
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row) R[col][row] = (col == row && col == 3) ? 1.0f : 0.0f;
    for (int col = 0; col < 3; ++col)
        for (int row = 0; row < 3; ++row)
            R[col][row] = (row == col ? c : 0.0f) + s * kx[row][col] + t * k[row] * k[col];
}

// out = A * B for matrices stored by columns
void matmul(const float A[4][4], const float B[4][4], float out[4][4])
{
    float r[4][4];
    for (int col = 0; col < 4; ++col)
        for (int row = 0; row < 4; ++row)
            r[col][row] = A[0][row] * B[col][0] + A[1][row] * B[col][1] + A[2][row] * B[col][2] + A[3][row] * B[col][3];
    std::memcpy(out, r, sizeof r);
}

void to_mat(const yn_camctl& c, float M[4][4])
{
    for (int i = 0; i < 4; ++i) {
        M[i][0] = c.view2world[i].x; M[i][1] = c.view2world[i].y;
        M[i][2] = c.view2world[i].z; M[i][3] = c.view2world[i].w;
    }
}

void from_mat(yn_camctl& c, const float M[4][4])
{
    for (int i = 0; i < 4; ++i) c.view2world[i] = {M[i][0], M[i][1], M[i][2], M[i][3]};
}

// Camera::setViewMatrix (src/Camera.cpp:105-117)
void set_view(yn_camctl& c, yn_float4 side, yn_float4 up, yn_float4 look, yn_float4 eye)
{
    cam_norm3(side); cam_norm3(up); cam_norm3(look);
    side.w = up.w = look.w = 0.0f;
    eye.w = 1.0f;
    c.side = side; c.up = up; c.look_at = look; c.eye = eye;
    c.view2world[0] = side;
    c.view2world[1] = up;
    c.view2world[2] = {-look.x, -look.y, -look.z, 0.0f};
    c.view2world[3] = eye;
}

}  // namespace

int yn_camctl_init(float y_fov, float rotation_speed, float move_speed, yn_camctl* out)
{
    if (!out || !(y_fov > 0.0f && y_fov < 180.0f)) { set_global_error("bad camera arguments"); return YN_ERR_INVALID; }
    std::memset(out, 0, sizeof *out);
    out->y_fov = y_fov;
    out->rotation_speed = rotation_speed;
    out->move_speed = move_speed;
    out->view_plane_dist = (float)(1.0 / std::tan((double)y_fov * 3.14 / 360.0));   // Camera.cpp:49
    set_view(*out, {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 0, 1});
    return 0;
}

int yn_camctl_set_view(yn_camctl* c, const float side[3], const float up[3], const float look_at[3], const float eye[3])
{
    if (!c || !side || !up || !look_at || !eye) { set_global_error("bad camera arguments"); return YN_ERR_INVALID; }
    set_view(*c, {side[0], side[1], side[2], 0}, {up[0], up[1], up[2], 0}, {look_at[0], look_at[1], look_at[2], 0},
             {eye[0], eye[1], eye[2], 1});
    return 0;
}

// Camera::setOrientation (src/Camera.cpp:119-166).  The reference's test
// `dir.length() > 0` is glm's component count (always 4), so the move step
// is always taken; with no movement and no rotation only the eye column is
// refreshed.
int yn_camctl_orient(yn_camctl* c, const float dir[3], float pitch, float yaw)
{
    if (!c || !dir) { set_global_error("bad camera arguments"); return YN_ERR_INVALID; }
    auto step = [&](const yn_float4& v, float sgn) {
        c->eye.x += sgn * v.x * c->move_speed;
        c->eye.y += sgn * v.y * c->move_speed;
        c->eye.z += sgn * v.z * c->move_speed;
    };
    if (dir[2] > 0) step(c->look_at, 1);
    else if (dir[2] < 0) step(c->look_at, -1);
    else if (dir[0] > 0) step(c->side, 1);
    else if (dir[0] < 0) step(c->side, -1);
    else if (dir[1] > 0) step(c->up, 1);
    else if (dir[1] < 0) step(c->up, -1);
    if (pitch == 0.0f && yaw == 0.0f) {
        c->view2world[3] = c->eye;
        return 0;
    }
    float rx[4][4], ry[4][4], M[4][4];
    rotation(pitch * c->rotation_speed, c->side.x, c->side.y, c->side.z, rx);
    rotation(yaw * c->rotation_speed, 0.0f, 1.0f, 0.0f, ry);
    // rotated up: is the camera still upright after the pitch?
    float up_y = rx[0][1] * c->up.x + rx[1][1] * c->up.y + rx[2][1] * c->up.z + rx[3][1] * c->up.w;
    to_mat(*c, M);
    M[3][0] = 0.0f; M[3][1] = 0.0f; M[3][2] = 0.0f; M[3][3] = 1.0f;      // to the origin
    if (up_y >= 0.0f) matmul(rx, M, M);
    matmul(ry, M, M);
    M[3][0] = c->eye.x; M[3][1] = c->eye.y; M[3][2] = c->eye.z; M[3][3] = c->eye.w;   // back to the eye
    from_mat(*c, M);
    c->side = c->view2world[0]; cam_norm3(c->side);
    c->up = c->view2world[1]; cam_norm3(c->up);
    yn_float4 l = c->view2world[2];
    cam_norm3(l);
    c->look_at = {-l.x, -l.y, -l.z, 0.0f};
    return 0;
}

// Camera::setBuffer (src/Camera.cpp:66-82): rows of view2world
int yn_camctl_pack(const yn_camctl* c, yn_cam* out)
{
    if (!c || !out) { set_global_error("bad camera arguments"); return YN_ERR_INVALID; }
    std::memset(out, 0, sizeof *out);
    const yn_float4* m = c->view2world;
    out->r1 = {m[0].x, m[1].x, m[2].x, m[3].x};
    out->r2 = {m[0].y, m[1].y, m[2].y, m[3].y};
    out->r3 = {m[0].z, m[1].z, m[2].z, m[3].z};
    out->r4 = {m[0].w, m[1].w, m[2].w, m[3].w};
    out->view_plane_dist = c->view_plane_dist;
    return 0;
}

}  // extern "C"
