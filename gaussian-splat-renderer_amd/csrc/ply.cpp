// Loader for the INRIA 3DGS .ply format (the reference's loadGaussianPly,
// scene/ply_loader.h:1-12 and 86-95, scene/ply_loader.cpp:100-150 for the decode): the file's
// raw fields become the activated f32 SoA scene that gsr_scene_create takes.
//
// Accepted: binary_little_endian with one list-free `vertex` element whose properties are
// found by name (any order, any scalar type; unknown properties are skipped but count in the
// record stride), SH degree 0..3 from the number of f_rest_* properties (0/9/24/45), other
// list-free elements before or after `vertex`. Rejected with GSR_ERROR_FORMAT and a message
// naming the problem: ASCII and big-endian files, PlayCanvas compressed PLY, missing required
// properties, f_rest counts other than 0/9/24/45 or with gaps, list properties in `vertex`,
// an empty `vertex`, files shorter than the header says.
//
// Activations (INRIA gaussian_model.py): alpha = 1 / (1 + exp(-opacity)), sigma = exp(log
// scale), q = the quaternion (w first) divided by its length, or the identity when its squared
// length is <= 1e-24 or not finite. SH: the file's f_dc_c and channel-major
// f_rest[c * (K - 1) + (k - 1)] become coefficient-major RGB triplets sh[(i * K + k) * 3 + c].
// Decoding runs on the host threads, one contiguous range of gaussians per thread.
#include "../../include/gsr.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <string_view>
#include <thread>
#include <vector>

namespace gsr {
gsr_status setError(gsr_status s, const char* message);  // capi.hip: the gsr_last_error text
}

namespace {

using gsr::setError;

enum class Scalar : uint8_t { I8, U8, I16, U16, I32, U32, F32, F64 };

uint32_t scalarBytes(Scalar t) {
    switch (t) {
        case Scalar::I8:
        case Scalar::U8: return 1;
        case Scalar::I16:
        case Scalar::U16: return 2;
        case Scalar::I32:
        case Scalar::U32:
        case Scalar::F32: return 4;
        case Scalar::F64: return 8;
    }
    return 0;
}

bool scalarFromName(std::string_view s, Scalar& t) {
    static const struct {
        const char* name;
        Scalar type;
    } names[] = {{"char", Scalar::I8},     {"int8", Scalar::I8},      {"uchar", Scalar::U8},   {"uint8", Scalar::U8},
                 {"short", Scalar::I16},   {"int16", Scalar::I16},    {"ushort", Scalar::U16}, {"uint16", Scalar::U16},
                 {"int", Scalar::I32},     {"int32", Scalar::I32},    {"uint", Scalar::U32},   {"uint32", Scalar::U32},
                 {"float", Scalar::F32},   {"float32", Scalar::F32},  {"double", Scalar::F64}, {"float64", Scalar::F64}};
    for (const auto& n : names)
        if (s == n.name) {
            t = n.type;
            return true;
        }
    return false;
}

struct Property {
    std::string name;
    Scalar type = Scalar::F32;
    bool list = false;
    uint32_t offset = 0;
};

struct Element {
    std::string name;
    uint64_t count = 0;
    uint32_t stride = 0;
    bool hasList = false;
    std::vector<Property> props;
    const Property* find(std::string_view n) const {
        for (const Property& p : props)
            if (p.name == n) return &p;
        return nullptr;
    }
};

struct Header {
    enum Format { LittleEndian, BigEndian, Ascii } format = LittleEndian;
    std::vector<Element> elements;
    std::vector<std::string> comments;
    size_t bytes = 0;
};

struct LoadError {
    std::string message;
};

std::vector<std::string_view> words(std::string_view line) {
    std::vector<std::string_view> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) ++i;
        const size_t b = i;
        while (i < line.size() && line[i] != ' ' && line[i] != '\t') ++i;
        if (i > b) out.push_back(line.substr(b, i - b));
    }
    return out;
}

bool parseUnsigned(std::string_view s, uint64_t& v) {
    if (s.empty() || s.size() > 19) return false;
    v = 0;
    for (char c : s) {
        if (c < '0' || c > '9') return false;
        v = v * 10 + uint64_t(c - '0');
    }
    return true;
}

constexpr size_t kMaxHeader = size_t(1) << 20;  // real 3DGS headers are ~1.5 KB

Header parseHeader(std::string_view bytes) {
    Header h;
    size_t pos = 0, line = 0;
    bool sawFormat = false;
    const std::string_view head = bytes.substr(0, std::min(bytes.size(), kMaxHeader));
    for (;;) {
        const size_t eol = head.find('\n', pos);
        if (eol == std::string_view::npos)
            throw LoadError{line == 0 ? "not a PLY file (empty or no header)"
                                      : "PLY header has no 'end_header' line (file truncated or not a PLY)"};
        std::string_view text = head.substr(pos, eol - pos);
        if (!text.empty() && text.back() == '\r') text.remove_suffix(1);  // CRLF headers
        pos = eol + 1;
        ++line;
        const std::string at = "PLY header line " + std::to_string(line) + ": ";
        if (line == 1) {
            if (text != "ply") throw LoadError{"not a PLY file (first line is not 'ply')"};
            continue;
        }
        const std::vector<std::string_view> w = words(text);
        if (w.empty()) continue;
        if (w[0] == "end_header") {
            if (!sawFormat) throw LoadError{"PLY header has no 'format' line"};
            h.bytes = pos;
            return h;
        }
        if (w[0] == "comment" || w[0] == "obj_info") {
            h.comments.emplace_back(text.size() > w[0].size() + 1 ? text.substr(w[0].size() + 1) : "");
        } else if (w[0] == "format") {
            if (w.size() < 2) throw LoadError{at + "malformed 'format' line"};
            if (w[1] == "binary_little_endian") h.format = Header::LittleEndian;
            else if (w[1] == "binary_big_endian") h.format = Header::BigEndian;
            else if (w[1] == "ascii") h.format = Header::Ascii;
            else throw LoadError{at + "unknown format '" + std::string(w[1]) + "'"};
            sawFormat = true;
        } else if (w[0] == "element") {
            Element e;
            if (w.size() != 3) throw LoadError{at + "malformed 'element' line"};
            e.name = std::string(w[1]);
            if (!parseUnsigned(w[2], e.count)) throw LoadError{at + "bad element count '" + std::string(w[2]) + "'"};
            h.elements.push_back(std::move(e));
        } else if (w[0] == "property") {
            if (h.elements.empty()) throw LoadError{at + "'property' before any 'element'"};
            Element& e = h.elements.back();
            Property p;
            if (w.size() == 5 && w[1] == "list") {
                Scalar a, b;
                if (!scalarFromName(w[2], a) || !scalarFromName(w[3], b)) throw LoadError{at + "unknown list types"};
                p.list = true;
                p.name = std::string(w[4]);
                e.hasList = true;
            } else if (w.size() == 3) {
                if (!scalarFromName(w[1], p.type))
                    throw LoadError{at + "unknown property type '" + std::string(w[1]) + "'"};
                p.name = std::string(w[2]);
                p.offset = e.stride;
                e.stride += scalarBytes(p.type);
            } else {
                throw LoadError{at + "malformed 'property' line"};
            }
            e.props.push_back(std::move(p));
        } else {
            throw LoadError{at + "unknown keyword '" + std::string(w[0]) + "'"};
        }
    }
}

struct Field {
    uint32_t offset = 0;
    Scalar type = Scalar::F32;
};

struct Layout {
    Field pos[3], dc[3], opacity, scale[3], rot[4];
    std::vector<Field> rest;  // f_rest_0 .. f_rest_{n-1}
    uint32_t stride = 0;
    bool allFloat = true;
    int degree = 0;
    uint64_t count = 0;
    uint64_t offset = 0;  // file offset of the first vertex record
    uint64_t after = 0;   // bytes of the list-free elements after `vertex`
    bool afterUnknown = false;
};

Layout planVertex(const Header& h, const std::string& file, uint64_t fileSize, std::vector<std::string>& warnings) {
    if (h.format == Header::Ascii)
        throw LoadError{file + ": ASCII PLY is not supported (3DGS files are binary_little_endian; re-export as binary)"};
    if (h.format == Header::BigEndian)
        throw LoadError{file + ": binary_big_endian PLY is not supported (expected binary_little_endian)"};
    for (const Element& e : h.elements) {
        bool packed = e.name == "chunk";
        for (const Property& p : e.props) packed = packed || p.name.rfind("packed_", 0) == 0;
        if (packed)
            throw LoadError{file + ": this is a PlayCanvas compressed PLY (element 'chunk' / packed_* properties), "
                                   "not a standard 3DGS PLY; decompress it first"};
    }
    Layout L;
    const Element* vertex = nullptr;
    L.offset = h.bytes;
    for (const Element& e : h.elements) {
        if (e.name == "vertex" && !vertex) {
            vertex = &e;
            continue;
        }
        if (!vertex) {
            if (e.hasList) throw LoadError{file + ": element '" + e.name + "' with list properties precedes 'vertex'"};
            const uint64_t left = fileSize > L.offset ? fileSize - L.offset : 0;
            if (e.stride != 0 && e.count > left / e.stride)
                throw LoadError{file + ": file is truncated inside element '" + e.name + "'"};
            L.offset += e.count * e.stride;
        } else if (e.hasList) {
            L.afterUnknown = true;
        } else {
            L.after += e.count * e.stride;
        }
        warnings.push_back("skipped element '" + e.name + "' (" + std::to_string(e.count) + " records)");
    }
    if (!vertex) throw LoadError{file + ": no 'vertex' element"};
    if (vertex->hasList) throw LoadError{file + ": 'vertex' element has list properties (not a 3DGS file)"};
    if (vertex->count == 0) throw LoadError{file + ": 'vertex' element is empty (0 gaussians)"};
    if (vertex->count > 0xFFFFFFFFull) throw LoadError{file + ": too many gaussians"};
    L.count = vertex->count;
    L.stride = vertex->stride;
    auto need = [&](const char* name) {
        const Property* p = vertex->find(name);
        if (!p) throw LoadError{file + ": missing required property '" + std::string(name) + "'"};
        L.allFloat = L.allFloat && p->type == Scalar::F32;
        return Field{p->offset, p->type};
    };
    static const char* posNames[3] = {"x", "y", "z"};
    static const char* dcNames[3] = {"f_dc_0", "f_dc_1", "f_dc_2"};
    static const char* scaleNames[3] = {"scale_0", "scale_1", "scale_2"};
    static const char* rotNames[4] = {"rot_0", "rot_1", "rot_2", "rot_3"};
    for (int i = 0; i < 3; ++i) L.pos[i] = need(posNames[i]);
    for (int i = 0; i < 3; ++i) L.dc[i] = need(dcNames[i]);
    L.opacity = need("opacity");
    for (int i = 0; i < 3; ++i) L.scale[i] = need(scaleNames[i]);
    for (int i = 0; i < 4; ++i) L.rot[i] = need(rotNames[i]);

    std::vector<const Property*> rest;
    for (const Property& p : vertex->props) {
        if (p.name.rfind("f_rest_", 0) != 0) continue;
        uint64_t idx = 0;
        if (!parseUnsigned(std::string_view(p.name).substr(7), idx) || idx > 4096)
            throw LoadError{file + ": bad SH property name '" + p.name + "'"};
        if (idx >= rest.size()) rest.resize(idx + 1, nullptr);
        if (rest[idx]) throw LoadError{file + ": duplicate property '" + p.name + "'"};
        rest[idx] = &p;
    }
    size_t present = 0;
    for (const Property* p : rest) present += p ? 1 : 0;
    if (present != rest.size())
        throw LoadError{file + ": f_rest_* indices are not contiguous (highest index " + std::to_string(rest.size() - 1) +
                        ", but only " + std::to_string(present) + " present)"};
    switch (present) {
        case 0: L.degree = 0; break;
        case 9: L.degree = 1; break;
        case 24: L.degree = 2; break;
        case 45: L.degree = 3; break;
        default:
            throw LoadError{file + ": unexpected number of f_rest_* properties: " + std::to_string(present) +
                            " (expected 0, 9, 24 or 45 for SH degree 0..3)"};
    }
    for (const Property* p : rest) {
        L.rest.push_back({p->offset, p->type});
        L.allFloat = L.allFloat && p->type == Scalar::F32;
    }
    // The payload must hold every record the header promises.
    const uint64_t avail = fileSize > L.offset ? fileSize - L.offset : 0;
    const uint64_t vertexBytes = L.count * L.stride;
    if (avail < vertexBytes)
        throw LoadError{file + ": file is truncated: header declares " + std::to_string(L.count) + " gaussians x " +
                        std::to_string(L.stride) + " bytes = " + std::to_string(vertexBytes) +
                        " bytes of vertex data, but only " + std::to_string(avail) + " bytes follow the header (" +
                        std::to_string(L.stride ? avail / L.stride : 0) + " complete records)"};
    if (!L.afterUnknown && avail > vertexBytes + L.after)
        warnings.push_back(std::to_string(avail - vertexBytes - L.after) +
                           " unexpected trailing bytes after the last element");
    if (!L.afterUnknown && avail < vertexBytes + L.after)
        warnings.push_back("elements after 'vertex' are truncated (ignored)");
    return L;
}

template <typename T>
T loadAs(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));  // records are packed: no alignment
    return v;
}

float toFloat(const uint8_t* p, Scalar t) {
    switch (t) {
        case Scalar::I8: return float(loadAs<int8_t>(p));
        case Scalar::U8: return float(loadAs<uint8_t>(p));
        case Scalar::I16: return float(loadAs<int16_t>(p));
        case Scalar::U16: return float(loadAs<uint16_t>(p));
        case Scalar::I32: return float(loadAs<int32_t>(p));
        case Scalar::U32: return float(loadAs<uint32_t>(p));
        case Scalar::F32: return loadAs<float>(p);
        case Scalar::F64: return float(loadAs<double>(p));
    }
    return 0.0f;
}

bool nonFinite(float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return (u & 0x7F800000u) == 0x7F800000u;
}

struct Scene {
    gsr_ply_scene pub{};
    std::vector<float> posAlpha, scales, rotations, sh, rawOpacity, rawLogScales, rawRotations;
    std::vector<std::string> warningText, commentText;
    std::vector<const char*> warningPtrs, commentPtrs;
};

template <bool kAllFloat>
void decode(const Layout& L, const uint8_t* base, size_t begin, size_t end, Scene& s, uint64_t& bad, uint32_t& zeroQ) {
    auto rd = [&](const uint8_t* rec, const Field& f) {
        const float v = kAllFloat ? loadAs<float>(rec + f.offset) : toFloat(rec + f.offset, f.type);
        bad += nonFinite(v) ? 1u : 0u;
        return v;
    };
    const uint32_t K = uint32_t((L.degree + 1) * (L.degree + 1));
    for (size_t i = begin; i < end; ++i) {
        const uint8_t* rec = base + i * L.stride;
        const float x = rd(rec, L.pos[0]), y = rd(rec, L.pos[1]), z = rd(rec, L.pos[2]);
        const float op = rd(rec, L.opacity);
        float ls[3], q[4];
        for (int a = 0; a < 3; ++a) ls[a] = rd(rec, L.scale[a]);
        for (int a = 0; a < 4; ++a) q[a] = rd(rec, L.rot[a]);
        s.rawOpacity[i] = op;
        for (int a = 0; a < 3; ++a) s.rawLogScales[3 * i + a] = ls[a];
        for (int a = 0; a < 4; ++a) s.rawRotations[4 * i + a] = q[a];
        float* pa = &s.posAlpha[4 * i];
        pa[0] = x;
        pa[1] = y;
        pa[2] = z;
        pa[3] = 1.0f / (1.0f + std::exp(-op));
        float* sc = &s.scales[4 * i];
        for (int a = 0; a < 3; ++a) sc[a] = std::exp(ls[a]);
        sc[3] = 0.0f;
        float* ro = &s.rotations[4 * i];
        const float n2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        if (n2 > 1e-24f && !nonFinite(n2)) {
            const float len = std::sqrt(n2);
            for (int a = 0; a < 4; ++a) ro[a] = q[a] / len;
        } else {
            ro[0] = 1.0f;  // identity, w first
            ro[1] = ro[2] = ro[3] = 0.0f;
            ++zeroQ;
        }
        float* out = &s.sh[size_t(i) * K * 3];
        for (uint32_t c = 0; c < 3; ++c) out[c] = rd(rec, L.dc[c]);
        for (uint32_t k = 1; k < K; ++k)
            for (uint32_t c = 0; c < 3; ++c) out[k * 3 + c] = rd(rec, L.rest[c * (K - 1) + (k - 1)]);
    }
}

gsr_status load(const uint8_t* bytes, uint64_t size, const std::string& file, gsr_ply_scene** out) {
    Scene* s = new (std::nothrow) Scene();
    if (!s) return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
    try {
        Header h;
        try {
            h = parseHeader(std::string_view(reinterpret_cast<const char*>(bytes), size_t(size)));
        } catch (const LoadError& e) {
            throw LoadError{file + ": " + e.message};
        }
        s->commentText = h.comments;
        const Layout L = planVertex(h, file, size, s->warningText);
        const size_t n = size_t(L.count);
        const uint32_t K = uint32_t((L.degree + 1) * (L.degree + 1));
        s->posAlpha.resize(4 * n);
        s->scales.resize(4 * n);
        s->rotations.resize(4 * n);
        s->sh.resize(n * K * 3);
        s->rawOpacity.resize(n);
        s->rawLogScales.resize(3 * n);
        s->rawRotations.resize(4 * n);

        const uint8_t* base = bytes + L.offset;
        const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
        const size_t threads = std::clamp<size_t>(n / 16384, 1, std::min<size_t>(hw, 64));
        std::vector<uint64_t> bad(threads, 0);
        std::vector<uint32_t> zeroQ(threads, 0);
        auto run = [&](size_t t) {
            const size_t b = n * t / threads, e = n * (t + 1) / threads;
            if (L.allFloat) decode<true>(L, base, b, e, *s, bad[t], zeroQ[t]);
            else decode<false>(L, base, b, e, *s, bad[t], zeroQ[t]);
        };
        if (threads == 1) {
            run(0);
        } else {
            std::vector<std::thread> pool;
            for (size_t t = 0; t < threads; ++t) pool.emplace_back(run, t);
            for (std::thread& th : pool) th.join();
        }
        gsr_ply_scene& p = s->pub;
        for (size_t t = 0; t < threads; ++t) {
            p.non_finite_values += bad[t];
            p.zero_quaternions += zeroQ[t];
        }
        if (p.non_finite_values)
            s->warningText.push_back(std::to_string(p.non_finite_values) + " NaN/Inf values in the file");
        if (p.zero_quaternions)
            s->warningText.push_back(std::to_string(p.zero_quaternions) + " zero-length quaternions replaced by identity");
        if (!L.allFloat) s->warningText.push_back("non-float32 properties converted to float");
        for (const std::string& w : s->warningText) s->warningPtrs.push_back(w.c_str());
        for (const std::string& c : s->commentText) s->commentPtrs.push_back(c.c_str());
        p.count = uint32_t(n);
        p.sh_degree = L.degree;
        p.pos_alpha = s->posAlpha.data();
        p.scales = s->scales.data();
        p.rotations = s->rotations.data();
        p.sh = s->sh.data();
        p.raw_opacity = s->rawOpacity.data();
        p.raw_log_scales = s->rawLogScales.data();
        p.raw_rotations = s->rawRotations.data();
        p.warning_count = uint32_t(s->warningPtrs.size());
        p.warnings = s->warningPtrs.data();
        p.comment_count = uint32_t(s->commentPtrs.size());
        p.comments = s->commentPtrs.data();
        p.internal = s;
    } catch (const LoadError& e) {
        delete s;
        return setError(GSR_ERROR_FORMAT, e.message.c_str());
    } catch (const std::bad_alloc&) {
        delete s;
        return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
    }
    *out = &s->pub;
    return GSR_OK;
}

std::string baseName(const char* path) {
    const std::string p(path);
    const size_t slash = p.find_last_of("/\\");
    return slash == std::string::npos ? p : p.substr(slash + 1);
}

}  // namespace

extern "C" {

gsr_status gsr_ply_load_memory(const void* bytes, uint64_t size, const char* name, gsr_ply_scene** out) {
    if (!out || (!bytes && size)) return setError(GSR_ERROR_INVALID_ARGUMENT, "gsr_ply_load_memory: null argument");
    *out = nullptr;
    return load(static_cast<const uint8_t*>(bytes), size, name ? std::string(name) : std::string("<memory>"), out);
}

gsr_status gsr_ply_load(const char* path, gsr_ply_scene** out) {
    if (!path || !out) return setError(GSR_ERROR_INVALID_ARGUMENT, "gsr_ply_load: null argument");
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return setError(GSR_ERROR_IO, (std::string("gsr_ply_load: cannot open ") + path).c_str());
    std::vector<uint8_t> bytes;
    try {
        if (std::fseek(f, 0, SEEK_END) != 0) throw 0;
        const long size = std::ftell(f);
        if (size < 0 || std::fseek(f, 0, SEEK_SET) != 0) throw 0;
        bytes.resize(size_t(size));
        if (size > 0 && std::fread(bytes.data(), 1, bytes.size(), f) != bytes.size()) throw 0;
    } catch (int) {
        std::fclose(f);
        return setError(GSR_ERROR_IO, (std::string("gsr_ply_load: cannot read ") + path).c_str());
    } catch (const std::bad_alloc&) {
        std::fclose(f);
        return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
    }
    std::fclose(f);
    return load(bytes.data(), bytes.size(), baseName(path), out);
}

void gsr_ply_free(gsr_ply_scene* scene) {
    if (scene) delete static_cast<Scene*>(scene->internal);
}

}  // extern "C"
