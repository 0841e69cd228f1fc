// scene_io.cpp — host-side scene sources: the reference's OBJ/MTL reader
// (src/Scene.cpp:62-410), newMaterial() defaults (include/CL_headers.h:
// 109-124), the camera packing of src/Camera.cpp:66-117, and the procedural
// scenes the benchmarks render (a 36-triangle Cornell box, rooms of
// icospheres).
#include "yn_internal.h"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>

namespace yune {

yn_material new_material()
{
    yn_material m;
    std::memset(&m, 0, sizeof m);
    m.ke = {0, 0, 0, 1};
    m.kd = {0.3f, 0.3f, 0.3f, 1};
    m.ks = {0, 0, 0, 1};
    m.n = 1; m.k = 1; m.px = 1; m.py = 1;
    m.alpha_x = 100; m.alpha_y = 100;
    m.is_specular = 0; m.is_transmissive = 0;
    return m;
}

namespace {

bool skip_line(const std::string& line) { return line.empty() || line[0] == '#'; }

// Scene::getMatFileName (Scene.cpp:385-410): `mtllib` must be the first
// non-comment line, otherwise there is none.
std::string mat_file_name(const std::string& obj)
{
    std::ifstream f(obj);
    std::string line, word;
    while (std::getline(f, line)) {
        if (skip_line(line)) continue;
        std::stringstream ss(line);
        ss >> word;
        if (word != "mtllib") return "";
        ss >> word;
        return word;
    }
    return "";
}

float read_float(std::stringstream& ss)
{
    std::string tok;
    if (!(ss >> tok)) return 0.0f;
    return std::strtof(tok.c_str(), nullptr);
}

}  // namespace

// Scene::loadModel (Scene.cpp:133-383).  Differences, both for a headless
// library: a missing `mtllib` gives the single default material without
// writing a default .mtl next to the model (Scene.cpp:160-176), and input
// the reference reads out of bounds on (a property before any `newmtl`,
// indices past the vertex list) is an error instead.
int load_obj(const std::string& obj_path, HostScene& out, std::string& err)
{
    out = HostScene();
    std::map<std::string, int> mat_index;
    std::string mat_fn = mat_file_name(obj_path);
    if (mat_fn.empty()) {
        mat_index["default"] = 0;
        out.mats.push_back(new_material());
    } else {
        std::string dir = obj_path.substr(0, obj_path.find_last_of('/') + 1);
        std::ifstream mf(dir + mat_fn);
        if (!mf.is_open()) { err = "Error opening material file."; return -5; }
        std::string line, word;
        while (std::getline(mf, line)) {
            if (skip_line(line)) continue;
            std::stringstream ss(line);
            ss >> word;
            if (word == "newmtl") {
                std::string id;
                ss >> id;
                mat_index[id] = (int)out.mats.size();
                out.mats.push_back(new_material());
                continue;
            }
            bool prop = word == "ke" || word == "kd" || word == "ks" || word == "n" || word == "k" || word == "px" ||
                        word == "py" || word == "alpha_x" || word == "alpha_y" || word == "is_specular" ||
                        word == "is_transmissive";
            if (!prop) continue;
            if (out.mats.empty()) { err = "Bad Material file."; return -5; }
            yn_material& m = out.mats.back();
            if (word == "ke" || word == "kd" || word == "ks") {
                float x = read_float(ss), y = read_float(ss), z = read_float(ss);
                yn_float4 v = {x, y, z, 1.0f};
                (word == "ke" ? m.ke : word == "kd" ? m.kd : m.ks) = v;
            } else if (word == "n") m.n = read_float(ss);
            else if (word == "k") m.k = read_float(ss);
            else if (word == "px") m.py = read_float(ss);          // swapped on read (Scene.cpp:210-213)
            else if (word == "py") m.px = read_float(ss);
            else if (word == "alpha_x") m.alpha_x = read_float(ss);
            else if (word == "alpha_y") m.alpha_y = read_float(ss);
            else if (word == "is_specular") { int v = 0; ss >> v; m.is_specular = v; }
            else if (word == "is_transmissive") { int v = 0; ss >> v; m.is_transmissive = v; }
        }
        if (out.mats.empty()) { err = "Bad Material file."; return -5; }
    }

    std::ifstream f(obj_path);
    if (!f.is_open()) { err = "Error opening object file..."; return -3; }
    std::vector<float> verts, norms;
    std::string line, word, matID;
    int idx = -1;
    while (std::getline(f, line)) {
        if (skip_line(line)) continue;
        std::stringstream ss(line);
        ss >> word;
        if (word == "mtllib") continue;
        if (word == "o") { matID.clear(); }
        else if (word == "v") { for (int k = 0; k < 3; ++k) verts.push_back(read_float(ss)); }
        else if (word == "vn") { for (int k = 0; k < 3; ++k) norms.push_back(read_float(ss)); }
        else if (word == "usemtl") {
            ss >> matID;
            auto it = mat_index.find(matID);
            idx = it != mat_index.end() ? it->second : 0;
        } else if (word == "f") {
            if (idx < 0) {
                auto it = mat_index.find(matID);
                idx = it != mat_index.end() ? it->second : 0;
            }
            yn_triangle t;
            std::memset(&t, 0, sizeof t);
            t.matID = idx;
            yn_float4* vv[3] = {&t.v1, &t.v2, &t.v3};
            yn_float4* nn[3] = {&t.vn1, &t.vn2, &t.vn3};
            for (int i = 0; i < 3; ++i) {
                std::string tok;
                ss >> tok;
                std::stringstream ts(tok);
                std::string part;
                int j = 0;
                while (std::getline(ts, part, '/')) {
                    if (j == 1) { ++j; continue; }          // vt skipped
                    if (j == 0 || j == 2) {
                        long k = std::strtol(part.c_str(), nullptr, 10) - 1;
                        const std::vector<float>& src = j == 0 ? verts : norms;
                        if (k < 0 || 3 * (size_t)k + 2 >= src.size()) {
                            err = "face index out of range: " + line; return -5;
                        }
                        yn_float4 v = {src[3 * k], src[3 * k + 1], src[3 * k + 2], j == 0 ? 1.0f : 0.0f};
                        *(j == 0 ? vv[i] : nn[i]) = v;
                    }
                    ++j;
                }
            }
            out.tris.push_back(t);
        }
    }
    return 0;
}

void make_camera(const float side[3], const float up[3], const float look_at[3], const float eye[3], float yfov_deg,
                 yn_cam& c)
{
    // glm::normalize(vec4 w=0): v * (1 / sqrt(dot(v, v)))
    auto nrm = [](const float in[3], float o[3]) {
        float inv = 1.0f / sqrtf(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
        o[0] = in[0] * inv; o[1] = in[1] * inv; o[2] = in[2] * inv;
    };
    float s[3], u[3], l[3];
    nrm(side, s); nrm(up, u); nrm(look_at, l);
    // view2world = mat4(side, up, -look_at, eye) by columns; setBuffer
    // packs its rows (Camera.cpp:66-82, 105-117).
    c.r1 = {s[0], u[0], -l[0], eye[0]};
    c.r2 = {s[1], u[1], -l[1], eye[1]};
    c.r3 = {s[2], u[2], -l[2], eye[2]};
    c.r4 = {0.0f, 0.0f, 0.0f, 1.0f};
    c.view_plane_dist = (float)(1.0 / std::tan((double)yfov_deg * 3.14 / 360.0));   // Camera.cpp:49
    c.pad[0] = c.pad[1] = c.pad[2] = 0.0f;
}

namespace {

yn_material diffuse(float r, float g, float b)
{
    yn_material m = new_material();
    m.kd = {r, g, b, 1};
    m.ks = {0, 0, 0, 1};
    m.alpha_x = 0; m.alpha_y = 0;     // Oren-Nayar sigma = 0: Lambert
    return m;
}

void push_tri(std::vector<yn_triangle>& T, const float a[3], const float b[3], const float c[3], const float n[3], int mat)
{
    yn_triangle t;
    std::memset(&t, 0, sizeof t);
    t.v1 = {a[0], a[1], a[2], 1}; t.v2 = {b[0], b[1], b[2], 1}; t.v3 = {c[0], c[1], c[2], 1};
    t.vn1 = t.vn2 = t.vn3 = {n[0], n[1], n[2], 0};
    t.matID = mat;
    T.push_back(t);
}

// quad p0,p1,p2,p3 (counter-clockwise seen from the normal side)
void push_quad(std::vector<yn_triangle>& T, const float p0[3], const float p1[3], const float p2[3], const float p3[3],
               const float n[3], int mat)
{
    push_tri(T, p0, p1, p2, n, mat);
    push_tri(T, p0, p2, p3, n, mat);
}

// axis-aligned box (sx, sy, sz) rotated about +Y by `deg`, bottom face at
// y0, centred at (cx, cz): 12 triangles, outward normals.
void push_box(std::vector<yn_triangle>& T, float cx, float y0, float cz, float sx, float sy, float sz, float deg, int mat)
{
    float a = deg * 3.14159265f / 180.0f, ca = std::cos(a), sa = std::sin(a);
    auto P = [&](float x, float y, float z, float o[3]) {
        o[0] = cx + ca * x + sa * z;
        o[1] = y0 + y;
        o[2] = cz - sa * x + ca * z;
    };
    auto N = [&](float x, float y, float z, float o[3]) { o[0] = ca * x + sa * z; o[1] = y; o[2] = -sa * x + ca * z; };
    float hx = sx / 2, hz = sz / 2;
    float c[8][3];
    int k = 0;
    for (int yy = 0; yy < 2; ++yy)
        for (int zz = 0; zz < 2; ++zz)
            for (int xx = 0; xx < 2; ++xx) P(xx ? hx : -hx, yy ? sy : 0, zz ? hz : -hz, c[k++]);
    // index: (y*4 + z*2 + x)
    float n[3];
    N(0, 1, 0, n);  push_quad(T, c[4], c[6], c[7], c[5], n, mat);   // top
    N(0, -1, 0, n); push_quad(T, c[0], c[1], c[3], c[2], n, mat);   // bottom
    N(0, 0, 1, n);  push_quad(T, c[2], c[3], c[7], c[6], n, mat);   // +z
    N(0, 0, -1, n); push_quad(T, c[0], c[4], c[5], c[1], n, mat);   // -z
    N(1, 0, 0, n);  push_quad(T, c[1], c[5], c[7], c[3], n, mat);   // +x
    N(-1, 0, 0, n); push_quad(T, c[0], c[2], c[6], c[4], n, mat);   // -x
}

}  // namespace

// Cornell box, 36 triangles, in the reference camera's default frame (eye at
// the origin looking down -z, yFOV 60; Camera.cpp:93-103): walls x,y in
// [-1,1], z in [-4,-2] (open towards the camera), red left / green right
// (the reference fixture's cornellbox.mtl colours), white elsewhere, all
// Oren-Nayar sigma 0; a 0.5 x 0.5 quad light 0.01 under the ceiling
// emitting (17,12,4) downwards, whose two triangles sit 0.005 above the
// analytic quad; a white diffuse box and a modified-Phong box with
// px + py = 50, standing 0.001 above the floor.
void gen_cornell(HostScene& S)
{
    S = HostScene();
    S.mats.push_back(diffuse(0.75f, 0.75f, 0.75f));   // 0 white
    S.mats.push_back(diffuse(0.8f, 0.0f, 0.0f));      // 1 red
    S.mats.push_back(diffuse(0.0f, 0.8f, 0.0f));      // 2 green
    yn_material lightm = new_material();
    lightm.ke = {17, 12, 4, 1};
    lightm.kd = {0, 0, 0, 1};
    lightm.alpha_x = 0; lightm.alpha_y = 0;
    S.mats.push_back(lightm);                          // 3 light panel
    yn_material glossy = new_material();
    glossy.kd = {0.25f, 0.25f, 0.25f, 1};
    glossy.ks = {0.5f, 0.5f, 0.5f, 1};
    glossy.px = 25; glossy.py = 25;
    glossy.alpha_x = 0; glossy.alpha_y = 0;
    S.mats.push_back(glossy);                          // 4 glossy
    const float x0 = -1, x1 = 1, y0 = -1, y1 = 1, z0 = -4, z1 = -2;
    auto& T = S.tris;
    {
        float a[3] = {x0, y0, z1}, b[3] = {x1, y0, z1}, c[3] = {x1, y0, z0}, d[3] = {x0, y0, z0}, n[3] = {0, 1, 0};
        push_quad(T, a, b, c, d, n, 0);               // floor
    }
    {
        float a[3] = {x0, y1, z0}, b[3] = {x1, y1, z0}, c[3] = {x1, y1, z1}, d[3] = {x0, y1, z1}, n[3] = {0, -1, 0};
        push_quad(T, a, b, c, d, n, 0);               // ceiling
    }
    {
        float a[3] = {x0, y0, z0}, b[3] = {x1, y0, z0}, c[3] = {x1, y1, z0}, d[3] = {x0, y1, z0}, n[3] = {0, 0, 1};
        push_quad(T, a, b, c, d, n, 0);               // back
    }
    {
        float a[3] = {x0, y0, z1}, b[3] = {x0, y0, z0}, c[3] = {x0, y1, z0}, d[3] = {x0, y1, z1}, n[3] = {1, 0, 0};
        push_quad(T, a, b, c, d, n, 1);               // left, red
    }
    {
        float a[3] = {x1, y0, z0}, b[3] = {x1, y0, z1}, c[3] = {x1, y1, z1}, d[3] = {x1, y1, z0}, n[3] = {-1, 0, 0};
        push_quad(T, a, b, c, d, n, 2);               // right, green
    }
    const float ly = 0.99f, lx0 = -0.25f, lz0 = -3.25f, ls = 0.5f;
    {
        float py = ly + 0.005f;
        float a[3] = {lx0, py, lz0}, b[3] = {lx0 + ls, py, lz0}, c[3] = {lx0 + ls, py, lz0 + ls},
              d[3] = {lx0, py, lz0 + ls}, n[3] = {0, -1, 0};
        push_quad(T, a, b, c, d, n, 3);               // light panel triangles
    }
    // boxes stand 0.001 above the floor so no face is coplanar with it
    push_box(T, 0.35f, -0.999f, -2.85f, 0.6f, 0.6f, 0.6f, -17.0f, 0);   // short white box
    push_box(T, -0.35f, -0.999f, -3.45f, 0.6f, 1.2f, 0.6f, 20.0f, 4);   // tall glossy box
    S.light.pos = {lx0, ly, lz0, 1};
    S.light.normal = {0, -1, 0, 0};
    S.light.ke = {17, 12, 4, 0};
    S.light.edge_l = {ls, 0, 0, 0};
    S.light.edge_w = {0, 0, ls, 0};
    float side[3] = {1, 0, 0}, up[3] = {0, 1, 0}, look[3] = {0, 0, -1}, eye[3] = {0, 0, 0};
    make_camera(side, up, look, eye, 60.0f, S.cam);
}

namespace {

struct V3 { float x, y, z; };

V3 mid_norm(V3 a, V3 b)
{
    V3 m = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
    float inv = 1.0f / sqrtf(m.x * m.x + m.y * m.y + m.z * m.z);
    return {m.x * inv, m.y * inv, m.z * inv};
}

// Unit icosphere: 20 * 4^subdiv triangles, each as 3 unit vectors.
std::vector<V3> unit_icosphere(int subdiv)
{
    const float t = (1.0f + sqrtf(5.0f)) / 2.0f;
    V3 v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) {
        float inv = 1.0f / sqrtf(p.x * p.x + p.y * p.y + p.z * p.z);
        p = {p.x * inv, p.y * inv, p.z * inv};
    }
    const int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<V3> tri;
    for (auto& ff : f) { tri.push_back(v[ff[0]]); tri.push_back(v[ff[1]]); tri.push_back(v[ff[2]]); }
    for (int s = 0; s < subdiv; ++s) {
        std::vector<V3> nt;
        nt.reserve(tri.size() * 4);
        for (size_t i = 0; i < tri.size(); i += 3) {
            V3 a = tri[i], b = tri[i + 1], c = tri[i + 2];
            V3 ab = mid_norm(a, b), bc = mid_norm(b, c), ca = mid_norm(c, a);
            V3 q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
            nt.insert(nt.end(), q, q + 12);
        }
        tri.swap(nt);
    }
    return tri;
}

// uniform float in [0,1) from a 32-bit draw
inline float unit(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

}  // namespace

// Room of icospheres: a closed 10 x 10 x 10 room (12 triangles, white
// diffuse), a 2 x 2 quad light 0.01 under the ceiling (2 panel triangles
// 0.005 above it), and `n_spheres` icospheres of 20 * 4^subdiv triangles
// with centres and radii (U[0.3,1.2]) drawn from std::mt19937(seed), kept
// inside the room.  Sphere i is diffuse (albedo U[0.2,0.9] per channel),
// modified Phong (exponent log-uniform in [10,1000] split as px = py) or
// dielectric (IOR 1.5, r0 = 0.04) for i mod 3 = 0, 1, 2.  Camera at
// (0,0,4.5) looking down -z, yFOV 60.
void gen_icosphere_room(uint32_t seed, int n_spheres, int subdiv, HostScene& S)
{
    S = HostScene();
    std::mt19937 rng(seed);
    S.mats.push_back(diffuse(0.75f, 0.75f, 0.75f));   // 0 room
    yn_material lightm = new_material();
    lightm.ke = {40, 32, 24, 1};
    lightm.kd = {0, 0, 0, 1};
    lightm.alpha_x = 0; lightm.alpha_y = 0;
    S.mats.push_back(lightm);                          // 1 light panel
    const float R = 5.0f;
    auto& T = S.tris;
    {
        float c[8][3];
        for (int i = 0; i < 8; ++i) { c[i][0] = (i & 1) ? R : -R; c[i][1] = (i & 2) ? R : -R; c[i][2] = (i & 4) ? R : -R; }
        float n[3];
        n[0] = 0; n[1] = 1; n[2] = 0;  push_quad(T, c[0], c[4], c[5], c[1], n, 0);   // floor y=-R
        n[1] = -1;                     push_quad(T, c[2], c[3], c[7], c[6], n, 0);   // ceiling
        n[0] = 1; n[1] = 0;            push_quad(T, c[0], c[2], c[6], c[4], n, 0);   // x=-R
        n[0] = -1;                     push_quad(T, c[1], c[5], c[7], c[3], n, 0);   // x=+R
        n[0] = 0; n[2] = 1;            push_quad(T, c[0], c[1], c[3], c[2], n, 0);   // z=-R
        n[2] = -1;                     push_quad(T, c[4], c[6], c[7], c[5], n, 0);   // z=+R
    }
    const float ly = R - 0.01f, lx0 = -1.0f, lz0 = -1.0f, ls = 2.0f;
    {
        float py = ly + 0.005f;
        float a[3] = {lx0, py, lz0}, b[3] = {lx0 + ls, py, lz0}, c[3] = {lx0 + ls, py, lz0 + ls},
              d[3] = {lx0, py, lz0 + ls}, n[3] = {0, -1, 0};
        push_quad(T, a, b, c, d, n, 1);
    }
    S.light.pos = {lx0, ly, lz0, 1};
    S.light.normal = {0, -1, 0, 0};
    S.light.ke = {40, 32, 24, 0};
    S.light.edge_l = {ls, 0, 0, 0};
    S.light.edge_w = {0, 0, ls, 0};
    std::vector<V3> ico = unit_icosphere(subdiv);
    T.reserve(T.size() + ico.size() / 3 * (size_t)n_spheres);
    for (int s = 0; s < n_spheres; ++s) {
        float r = 0.3f + 0.9f * unit(rng());
        float lim = R - r - 0.01f;
        float cx = -lim + 2 * lim * unit(rng()), cy = -lim + 2 * lim * unit(rng()), cz = -lim + 2 * lim * unit(rng());
        yn_material m = new_material();
        m.alpha_x = 0; m.alpha_y = 0;
        if (s % 3 == 0) {
            m.kd = {0.2f + 0.7f * unit(rng()), 0.2f + 0.7f * unit(rng()), 0.2f + 0.7f * unit(rng()), 1};
            m.ks = {0, 0, 0, 1};
        } else if (s % 3 == 1) {
            float e = std::exp(std::log(10.0f) + (std::log(1000.0f) - std::log(10.0f)) * unit(rng()));
            float kd = 0.1f + 0.2f * unit(rng()), ks = 0.3f + 0.3f * unit(rng());
            m.kd = {kd, kd, kd, 1};
            m.ks = {ks, ks, ks, 1};
            m.px = e / 2; m.py = e / 2;
        } else {
            m.is_specular = 1; m.is_transmissive = 1;
            m.n = 1.5f;
            m.kd = {0, 0, 0, 1};
            m.ks = {0.04f, 0.04f, 0.04f, 1};
        }
        int mid = (int)S.mats.size();
        S.mats.push_back(m);
        for (size_t i = 0; i < ico.size(); i += 3) {
            float p[3][3], nn[3][3];
            for (int k = 0; k < 3; ++k) {
                const V3& u = ico[i + k];
                p[k][0] = cx + r * u.x; p[k][1] = cy + r * u.y; p[k][2] = cz + r * u.z;
                nn[k][0] = u.x; nn[k][1] = u.y; nn[k][2] = u.z;
            }
            yn_triangle t;
            std::memset(&t, 0, sizeof t);
            t.v1 = {p[0][0], p[0][1], p[0][2], 1}; t.v2 = {p[1][0], p[1][1], p[1][2], 1}; t.v3 = {p[2][0], p[2][1], p[2][2], 1};
            t.vn1 = {nn[0][0], nn[0][1], nn[0][2], 0}; t.vn2 = {nn[1][0], nn[1][1], nn[1][2], 0};
            t.vn3 = {nn[2][0], nn[2][1], nn[2][2], 0};
            t.matID = mid;
            T.push_back(t);
        }
    }
    float side[3] = {1, 0, 0}, up[3] = {0, 1, 0}, look[3] = {0, 0, -1}, eye[3] = {0, 0, 4.5f};
    make_camera(side, up, look, eye, 60.0f, S.cam);
}

}  // namespace yune
