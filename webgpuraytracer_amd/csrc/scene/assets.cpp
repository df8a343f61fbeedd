// assets.cpp — asset ingest for the host scene builder: Radiance RGBE
// (.hdr) images as lib/parse-hdr.js reads them (parse-hdr.js:20-211, used by
// HDRTexture._load, tracer.js:53-73) and Wavefront OBJ meshes as
// ModelData.parseOBJ and generateTriangles read them (tracer.js:949-1022).
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pt_scene.h"

namespace {

// ---------------------------------------------------------------- RGBE -----
struct HdrHeader {
    uint32_t width = 0, height = 0;
    double exposure = 1.0;
    bool rle = false;
    uint64_t data_offset = 0;
};

bool contains(const std::string& s, const char* pat) { return s.find(pat) != std::string::npos; }

// The first match of "EXPOSURE=\s*([0-9]*[.][0-9]*)" in the line: its number
// (NaN for a lone "."), or false when the pattern does not occur.
bool match_exposure(const std::string& s, double* v) {
    for (size_t at = s.find("EXPOSURE="); at != std::string::npos; at = s.find("EXPOSURE=", at + 1)) {
        size_t i = at + 9;
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\f' || s[i] == '\v')) ++i;
        const size_t b = i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        if (i >= s.size() || s[i] != '.') continue;
        ++i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        const std::string num = s.substr(b, i - b);
        *v = num == "." ? NAN : std::strtod(num.c_str(), nullptr);
        return true;
    }
    return false;
}

// The first match of "-Y ([0-9]+) \+X ([0-9]+)".
bool match_size(const std::string& s, uint32_t* h, uint32_t* w) {
    for (size_t at = s.find("-Y "); at != std::string::npos; at = s.find("-Y ", at + 1)) {
        size_t i = at + 3, b = i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        if (i == b || s.compare(i, 4, " +X ") != 0) continue;
        const std::string hs = s.substr(b, i - b);
        i += 4;
        const size_t c = i;
        while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        if (i == c) continue;
        const unsigned long long hv = std::strtoull(hs.c_str(), nullptr, 10);
        const unsigned long long wv = std::strtoull(s.substr(c, i - c).c_str(), nullptr, 10);
        if (hv > 0xffffffffull || wv > 0xffffffffull) return false;
        *h = uint32_t(hv);
        *w = uint32_t(wv);
        return true;
    }
    return false;
}

// parseHdr's header loop (parse-hdr.js:154-175): at most 20 lines, each
// tested against the patterns in the reference's order; the size line ends it.
int hdr_header(const uint8_t* b, uint64_t n, HdrHeader* h, std::string* err) {
    uint64_t off = 0;
    for (int line_no = 0; line_no < 20 && off < n; ++line_no) {
        std::string line;
        while (off < n && b[off] != '\n') line.push_back(char(b[off++]));
        if (off < n) ++off;  // the newline
        double e;
        if (contains(line, "#?RADIANCE")) {
        } else if (contains(line, "FORMAT=32-bit_rle_rgbe")) {
            h->rle = true;
        } else if (match_exposure(line, &e)) {
            h->exposure = e;
        } else if (contains(line, "#")) {
        } else if (match_size(line, &h->height, &h->width)) {
            break;
        }
    }
    h->data_offset = off;
    if (!h->rle) { *err = "File is not run length encoded!"; return PT_ERR_LAYOUT; }
    if (h->width == 0 || h->height == 0 || uint64_t(h->width) * h->height > (1ull << 28)) {
        *err = "missing or oversized '-Y H +X W' size line";
        return PT_ERR_LAYOUT;
    }
    return PT_OK;
}

// readPixelsRawRLE (parse-hdr.js:21-132): new-style run-length scanlines
// (2, 2, width hi, width lo, then each channel's runs and literals); a first
// scanline without that marker means the rest of the file is flat RGBE.
int hdr_pixels(const uint8_t* b, uint64_t n, uint64_t off, uint32_t width, uint32_t height, uint8_t* out,
               std::string* err) {
    uint64_t o = 0;  // output byte
    std::vector<uint8_t> line(size_t(width) * 4);
    for (uint32_t left = height; left > 0; --left) {
        if (off + 4 > n) { *err = "Error reading bytes: expected 4"; return PT_ERR_LAYOUT; }
        const uint8_t* rgbe = b + off;
        off += 4;
        if (rgbe[0] != 2 || rgbe[1] != 2 || (rgbe[2] & 0x80) != 0) {
            const uint64_t rest = uint64_t(width) * left * 4 - 4;
            if (off + rest > n) { *err = "Error reading raw pixels"; return PT_ERR_LAYOUT; }
            std::memcpy(out + o, rgbe, 4);
            std::memcpy(out + o + 4, b + off, rest);
            return PT_OK;
        }
        if ((uint32_t(rgbe[2]) << 8 | rgbe[3]) != width) { *err = "Wrong scanline width"; return PT_ERR_LAYOUT; }
        uint32_t ptr = 0;
        for (uint32_t ch = 0; ch < 4; ++ch) {
            const uint32_t end = (ch + 1) * width;
            while (ptr < end) {
                if (off + 2 > n) { *err = "Error reading 2-byte buffer"; return PT_ERR_LAYOUT; }
                uint32_t count = b[off];
                const uint8_t v = b[off + 1];
                off += 2;
                if (count > 128) {  // a run of one value
                    count -= 128;
                    if (count == 0 || count > end - ptr) { *err = "Bad scanline data"; return PT_ERR_LAYOUT; }
                    std::memset(&line[ptr], v, count);
                    ptr += count;
                } else {  // literals: the byte already read, then count - 1 more
                    if (count == 0 || count > end - ptr) { *err = "Bad scanline data"; return PT_ERR_LAYOUT; }
                    line[ptr++] = v;
                    if (--count > 0) {
                        if (off + count > n) { *err = "Error reading non-run data"; return PT_ERR_LAYOUT; }
                        std::memcpy(&line[ptr], b + off, count);
                        off += count;
                        ptr += count;
                    }
                }
            }
        }
        for (uint32_t i = 0; i < width; ++i, o += 4) {
            out[o] = line[i];
            out[o + 1] = line[i + width];
            out[o + 2] = line[i + 2 * width];
            out[o + 3] = line[i + 3 * width];
        }
    }
    return PT_OK;
}

// ----------------------------------------------------------------- OBJ -----
struct ObjCorner {
    long v, u, n;  // zero-based, -1 when absent
};
struct ObjData {
    std::vector<float> pos, uv, nrm;
    std::vector<ObjCorner> corners;  // 3 per triangle
};

// An OBJ index: 1-based, or negative = relative to the end of its list so
// far.  Returns -1 for an empty field.
long obj_index(const std::string& f, size_t count) {
    if (f.empty()) return -1;
    const long i = std::strtol(f.c_str(), nullptr, 10);
    if (i > 0) return i - 1;
    if (i < 0) return long(count) + i;
    return -1;
}

// parseOBJ (tracer.js:949-987): v, vt, vn and f lines, whitespace separated;
// faces as v, v/vt, v/vt/vn or v//vn, fan-triangulated around the first corner.
void parse_obj(const char* text, uint64_t n, ObjData* d) {
    uint64_t off = 0;
    std::vector<std::string> parts;
    while (off < n) {
        uint64_t e = off;
        while (e < n && text[e] != '\n') ++e;
        parts.clear();
        uint64_t i = off;
        while (i < e) {
            while (i < e && std::isspace(static_cast<unsigned char>(text[i]))) ++i;
            const uint64_t b = i;
            while (i < e && !std::isspace(static_cast<unsigned char>(text[i]))) ++i;
            if (i > b) parts.emplace_back(text + b, text + i);
        }
        off = e + 1;
        if (parts.empty()) continue;
        const std::string& t = parts[0];
        auto num = [&](size_t k) { return k < parts.size() ? float(std::strtod(parts[k].c_str(), nullptr)) : NAN; };
        if (t == "v") {
            for (size_t k = 1; k <= 3; ++k) d->pos.push_back(num(k));
        } else if (t == "vt") {
            for (size_t k = 1; k <= 2; ++k) d->uv.push_back(num(k));
        } else if (t == "vn") {
            for (size_t k = 1; k <= 3; ++k) d->nrm.push_back(num(k));
        } else if (t == "f" && parts.size() >= 4) {
            std::vector<ObjCorner> face;
            for (size_t k = 1; k < parts.size(); ++k) {
                const std::string& p = parts[k];
                const size_t s1 = p.find('/');
                const size_t s2 = s1 == std::string::npos ? std::string::npos : p.find('/', s1 + 1);
                const std::string fv = p.substr(0, s1);
                const std::string fu = s1 == std::string::npos ? "" : p.substr(s1 + 1, s2 == std::string::npos
                                                                                           ? std::string::npos
                                                                                           : s2 - s1 - 1);
                const std::string fn = s2 == std::string::npos ? "" : p.substr(s2 + 1);
                face.push_back(ObjCorner{obj_index(fv, d->pos.size() / 3), obj_index(fu, d->uv.size() / 2),
                                         obj_index(fn, d->nrm.size() / 3)});
            }
            for (size_t k = 1; k + 1 < face.size(); ++k) {
                d->corners.push_back(face[0]);
                d->corners.push_back(face[k]);
                d->corners.push_back(face[k + 1]);
            }
        }
    }
}

// generateTriangles (tracer.js:989-1022) with one vertex per corner.  A
// corner without a normal takes its face's geometric normal and one without
// texture coordinates takes (0, 0) (the reference reads undefined there,
// which the Float32Array stores as NaN).  Returns false on a position index
// outside the file's positions.
bool obj_corners(const ObjData& d, float* pos, float* nrm, float* uv) {
    const size_t np = d.pos.size() / 3, nn = d.nrm.size() / 3, nu = d.uv.size() / 2;
    for (size_t t = 0; t * 3 < d.corners.size(); ++t) {
        double p[3][3];
        for (int k = 0; k < 3; ++k) {
            const ObjCorner& c = d.corners[3 * t + k];
            if (c.v < 0 || size_t(c.v) >= np) return false;
            for (int a = 0; a < 3; ++a) p[k][a] = d.pos[3 * size_t(c.v) + a];
        }
        double fn[3] = {(p[1][1] - p[0][1]) * (p[2][2] - p[0][2]) - (p[1][2] - p[0][2]) * (p[2][1] - p[0][1]),
                        (p[1][2] - p[0][2]) * (p[2][0] - p[0][0]) - (p[1][0] - p[0][0]) * (p[2][2] - p[0][2]),
                        (p[1][0] - p[0][0]) * (p[2][1] - p[0][1]) - (p[1][1] - p[0][1]) * (p[2][0] - p[0][0])};
        const double len = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        for (int a = 0; a < 3; ++a) fn[a] = len > 0.0 ? fn[a] / len : (a == 1 ? 1.0 : 0.0);
        for (int k = 0; k < 3; ++k) {
            const ObjCorner& c = d.corners[3 * t + k];
            const size_t o = 3 * (3 * t + k);
            for (int a = 0; a < 3; ++a) {
                pos[o + a] = float(p[k][a]);
                nrm[o + a] = (c.n >= 0 && size_t(c.n) < nn) ? d.nrm[3 * size_t(c.n) + a] : float(fn[a]);
            }
            const size_t ou = 2 * (3 * t + k);
            const bool has_uv = c.u >= 0 && size_t(c.u) < nu;
            uv[ou] = has_uv ? d.uv[2 * size_t(c.u)] : 0.0f;
            uv[ou + 1] = has_uv ? d.uv[2 * size_t(c.u) + 1] : 0.0f;
        }
    }
    return true;
}

thread_local std::string g_asset_err;

int asset_fail(int rc, const std::string& msg) {
    g_asset_err = msg;
    return rc;
}

}  // namespace

extern "C" {

const char* ptgen_asset_error(void) { return g_asset_err.c_str(); }

int ptgen_parse_hdr(const uint8_t* bytes, uint64_t n, uint32_t* width, uint32_t* height, float* exposure,
                    float* rgba, uint64_t rgba_count) {
    if (!bytes || !width || !height) return asset_fail(PT_ERR_ARG, "NULL argument");
    HdrHeader h;
    std::string err;
    int rc = hdr_header(bytes, n, &h, &err);
    if (rc) return asset_fail(rc, err);
    *width = h.width;
    *height = h.height;
    if (exposure) *exposure = float(h.exposure);
    if (!rgba) return PT_OK;  // size query
    const uint64_t px = uint64_t(h.width) * h.height;
    if (rgba_count != px * 4) return asset_fail(PT_ERR_ARG, "rgba must hold width * height * 4 floats");
    std::vector<uint8_t> data(px * 4);
    if ((rc = hdr_pixels(bytes, n, h.data_offset, h.width, h.height, data.data(), &err))) return asset_fail(rc, err);
    // parse-hdr.js:188-205: channel / 255 * 2^(e - 128) in double, stored as float; alpha 1
    for (uint64_t i = 0; i < px; ++i) {
        const double f = std::pow(2.0, double(data[4 * i + 3]) - 128.0);
        for (int c = 0; c < 3; ++c) rgba[4 * i + c] = float(double(data[4 * i + c]) / 255.0 * f);
        rgba[4 * i + 3] = 1.0f;
    }
    return PT_OK;
}

int ptgen_obj_triangles(const char* text, uint64_t n, uint32_t* n_triangles) {
    if (!text || !n_triangles) return asset_fail(PT_ERR_ARG, "NULL argument");
    ObjData d;
    parse_obj(text, n, &d);
    if (d.corners.size() / 3 > 0xffffffffull) return asset_fail(PT_ERR_CAPACITY, "too many triangles");
    *n_triangles = uint32_t(d.corners.size() / 3);
    return PT_OK;
}

int ptgen_obj_read(const char* text, uint64_t n, float* positions, float* normals, float* uvs, uint32_t n_triangles) {
    if (!text || !positions || !normals || !uvs) return asset_fail(PT_ERR_ARG, "NULL argument");
    ObjData d;
    parse_obj(text, n, &d);
    if (d.corners.size() != size_t(n_triangles) * 3) return asset_fail(PT_ERR_ARG, "triangle count mismatch");
    if (!obj_corners(d, positions, normals, uvs)) return asset_fail(PT_ERR_LAYOUT, "face index out of range");
    return PT_OK;
}

int ptgen_add_obj(ptgen_scene* s, const char* text, uint64_t n) {
    if (!s || !text) return PT_ERR_ARG;
    ObjData d;
    parse_obj(text, n, &d);
    const size_t tris = d.corners.size() / 3;
    if (tris == 0) return asset_fail(PT_ERR_LAYOUT, "no faces");
    if (tris > (1u << 24)) return asset_fail(PT_ERR_CAPACITY, "too many triangles");
    std::vector<float> pos(tris * 9), nrm(tris * 9), uv(tris * 6);
    if (!obj_corners(d, pos.data(), nrm.data(), uv.data())) return asset_fail(PT_ERR_LAYOUT, "face index out of range");
    std::vector<uint32_t> idx(tris * 3);
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
    return ptgen_add_model(s, pos.data(), nrm.data(), uv.data(), uint32_t(tris * 3), idx.data(), uint32_t(tris));
}

}  // extern "C"
