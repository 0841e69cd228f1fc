// capi_host.cpp — host-only entry points of the C ABI: BVH build, scene
// loading / generation and camera packing (include/yune_hip.h).
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

}  // extern "C"
