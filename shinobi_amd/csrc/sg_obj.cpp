// sg_obj.cpp -- scene files on the host: Wavefront OBJ geometry and the
// reference's camera.cfg.
//
// OBJ follows FWObject::read (src/Loaders/FWObjectReader.cpp:181-502) for
// everything that decides the triangles: which lines are accepted, how
// numbers are converted ((float)strtod for coordinates, strtol for indices)
// and the fan triangulation of polygons.  Materials, normals and texture
// coordinates do not reach the grid and are only parsed as far as their
// syntax decides whether a line is accepted.  A line the reference rejects
// ("Error at line ...") is skipped and counted.  Two inputs on which the
// reference has undefined behaviour are handled here: a face index outside
// the vertex list (it would read past the array; here the load fails with
// SG_ERR_INVALID_SCENE) and a face of more than 4096 vertices (it would
// overrun its index buffer; here the line is rejected).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/shinobi_grid.h"

struct sg_obj {
    std::vector<float> v;          // xyz per vertex
    std::vector<long> tri;         // vertex indices, 3 per triangle
    sg_obj_info info{};
};

namespace {

void skip_ws(const char*& s) {
    while (*s && isspace((unsigned char)*s)) ++s;
}

constexpr size_t kMaxFaceVerts = 4096;

// one 'v' / 'vn' / 'vt' line after the 'v'; false = rejected line
bool parse_vertex(const char* s, sg_obj& o) {
    int type;   // 0 position, 1 normal, 2 texture coordinate
    const char c = (char)tolower((unsigned char)*s);
    if (c == 'n') {
        type = 1;
        ++s;
    } else if (c == 't') {
        type = 2;
        ++s;
    } else if (isspace((unsigned char)*s)) {
        type = 0;
    } else {
        return false;
    }
    float xyz[3];
    int k = 0;
    for (;;) {
        skip_ws(s);
        if (*s == 0) break;
        char* end;
        const float f = (float)strtod(s, &end);
        if (end == s) return false;
        s = end;
        if (k >= 3) return false;
        xyz[k++] = f;
    }
    if ((type != 2 && k != 3) || (type == 2 && k < 2)) return false;
    if (type == 0) o.v.insert(o.v.end(), xyz, xyz + 3);
    return true;
}

// one face line after the 'f' (and an optional 'o'); vertex references are
// v, v/t, v//n or v/t/n; false = rejected line
bool parse_face(const char* s, std::vector<long>& verts) {
    verts.clear();
    skip_ws(s);
    for (;;) {
        if (verts.size() >= kMaxFaceVerts) return false;
        char* end;
        long idx = strtol(s, &end, 10);
        if (end == s) return false;
        s = end;
        verts.push_back(idx - 1);
        skip_ws(s);
        if (*s == '/') {
            ++s;
            skip_ws(s);
            if (*s != '/') {
                strtol(s, &end, 10);   // texture coordinate index
                if (end == s) return false;
                s = end;
            }
            skip_ws(s);
            if (*s == '/') {
                ++s;
                skip_ws(s);
                idx = strtol(s, &end, 10);   // normal index
                // the reader takes the number only when what follows it is
                // another index or the end of the line (or a comment)
                const char* look = end;
                skip_ws(look);
                if (isdigit((unsigned char)*look) || ((*look == 0 || *look == '#') && end != s)) {
                    if (end == s) return false;
                    s = end;
                }
            }
        }
        skip_ws(s);
        if (*s == 0) break;
    }
    return verts.size() >= 3;
}

}  // namespace

extern "C" sg_status sg_obj_load(const char* path, sg_obj** out, sg_obj_info* info) {
    if (!path || !out) return SG_ERR_INVALID_ARG;
    *out = nullptr;
    std::ifstream in(path, std::ios::in);
    if (in.fail()) return SG_ERR_IO;
    sg_obj* o = new sg_obj();
    std::string buf;
    std::vector<long> verts;
    uint64_t bad = 0, faces = 0;
    bool index_error = false;
    while (std::getline(in, buf)) {
        const char* s = buf.c_str();
        skip_ws(s);
        bool ok = true;
        switch (tolower((unsigned char)*s)) {
            case 0:
            case 'o':
            case 'g':
            case 's':
            case '#':
                break;
            case 'v':
                ok = parse_vertex(s + 1, *o);
                break;
            case 'f': {
                ++s;
                if (tolower((unsigned char)*s) == 'o') ++s;
                ok = parse_face(s, verts);
                if (ok) {
                    ++faces;
                    for (size_t k = 1; k + 1 < verts.size(); ++k) {
                        o->tri.push_back(verts[0]);
                        o->tri.push_back(verts[k]);
                        o->tri.push_back(verts[k + 1]);
                    }
                }
                break;
            }
            default:
                // usemtl / mtllib / unknown entities: no geometry; a usemtl or
                // mtllib without a name is a rejected line
                if (strncasecmp(s, "usemtl", 6) == 0 || strncasecmp(s, "mtllib", 6) == 0) {
                    const char* n = s + 6;
                    skip_ws(n);
                    ok = *n != 0;
                }
                break;
        }
        if (!ok) ++bad;
    }
    if (in.bad()) {
        delete o;
        return SG_ERR_IO;
    }
    const long nv = (long)(o->v.size() / 3);
    for (long t : o->tri)
        if (t < 0 || t >= nv) index_error = true;
    if (index_error || o->tri.size() / 3 > 0xFFFFFFFFull) {
        delete o;
        return SG_ERR_INVALID_SCENE;
    }
    o->info.n_vertices = (uint32_t)nv;
    o->info.n_faces = (uint32_t)faces;
    o->info.n_triangles = (uint32_t)(o->tri.size() / 3);
    o->info.n_bad_lines = (uint32_t)bad;
    if (info) *info = o->info;
    *out = o;
    return SG_OK;
}

extern "C" sg_status sg_obj_triangles(const sg_obj* o, float* const soa9[9]) {
    if (!o || !soa9) return SG_ERR_INVALID_ARG;
    for (int k = 0; k < 9; ++k)
        if (!soa9[k] && o->info.n_triangles) return SG_ERR_INVALID_ARG;
    for (size_t i = 0; i < o->info.n_triangles; ++i)
        for (int c = 0; c < 3; ++c) {
            const float* p = &o->v[3 * (size_t)o->tri[3 * i + c]];
            soa9[3 * c + 0][i] = p[0];
            soa9[3 * c + 1][i] = p[1];
            soa9[3 * c + 2][i] = p[2];
        }
    return SG_OK;
}

extern "C" void sg_obj_destroy(sg_obj* o) { delete o; }

// CameraManager::read (src/Loaders/CameraManager.cpp:93-165): keyword lines
// parsed with stream extraction; every line that contains '#' is skipped.
// rotation is stored as read; the reference applies it to the frame and
// resets it, which for the zero rotation its files hold is the identity.
extern "C" sg_status sg_camera_load(const char* path, sg_camera* out) {
    if (!path || !out) return SG_ERR_INVALID_ARG;
    std::ifstream in(path);
    if (!in) return SG_ERR_IO;
    sg_camera c{};
    std::string line, key;
    while (std::getline(in, line)) {
        if (line.find('#') != std::string::npos) continue;
        std::stringstream ss(line);
        key.clear();
        ss >> key;
        float* v3 = nullptr;
        if (key == "position") v3 = c.position;
        else if (key == "orientation") v3 = c.orientation;
        else if (key == "up") v3 = c.up;
        else if (key == "rotation") v3 = c.rotation;
        if (v3) {
            ss >> v3[0] >> v3[1] >> v3[2];
        } else if (key == "resX") {
            int r = 0;
            ss >> r;
            c.res_x = r > 0 ? (uint32_t)r : 0u;
        } else if (key == "resY") {
            int r = 0;
            ss >> r;
            c.res_y = r > 0 ? (uint32_t)r : 0u;
        } else if (key == "FOV") {
            ss >> c.fov_deg;
        }
    }
    *out = c;
    return SG_OK;
}
