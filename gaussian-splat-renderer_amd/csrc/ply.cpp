// Loader for the INRIA 3DGS .ply format: the drop-in for the reference's loadGaussianPly
// (scene/ply_loader.h:1-12, 86-95). The file's raw fields become the activated f32 SoA scene that
// gsr_scene_create takes. Accepted and rejected files and the activations follow the reference's
// behaviour (scene/ply_loader.cpp:100-150); the messages name the same problems.
//
// Design (three stages, each a table or a plan rather than per-case code):
//   1. Header -> Schema. Every header line is split into words and dispatched by its keyword
//      through a table of handlers (ply, format, comment/obj_info, element, property,
//      end_header); the handlers build the list of elements with their properties, byte
//      offsets and record strides.
//   2. Schema -> ColumnPlan. The `vertex` element's properties are resolved by name into
//      gaussian attributes (position, f_dc, f_rest, opacity, scale, rotation). Each used property
//      becomes one column: its offset in the record, a reader for its scalar type (one entry of
//      the scalar table) and its destination, an output array with a per-gaussian stride.
//      Properties the scene does not use are skipped but count in the stride.
//   3. Decode. Host threads claim chunks of 8192 records from an atomic counter; a chunk is
//      decoded column by column: each column's values are read from the chunk's records and
//      stored straight into their final arrays (the raw opacity / log-scale / rotation arrays,
//      the positions, the coefficient-major SH array, whose f_rest columns carry their
//      transposed slot). An activation pass over the chunk's gaussians then fills alpha, sigma
//      and the normalised rotation from the raw arrays.
//
// Activations (INRIA gaussian_model.py): alpha = 1 / (1 + exp(-opacity)), sigma = exp(log
// scale), q = the quaternion (w first) divided by its length; a quaternion whose squared length
// is not above 1e-24 or is not finite becomes the identity (counted). NaN/Inf values in used
// properties are counted.
#include "../../include/gsr.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
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

// --- Scalar types ------------------------------------------------------------------------------

using Reader = float (*)(const uint8_t*);

template <typename T>
float readScalar(const uint8_t* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));  // records are packed: no alignment
    return static_cast<float>(v);
}

struct ScalarType {
    const char* names[2];  // the PLY name and its sized alias
    uint32_t bytes;
    Reader read;
};

constexpr std::array<ScalarType, 8> kScalars = {{
    {{"char", "int8"}, 1, readScalar<int8_t>},
    {{"uchar", "uint8"}, 1, readScalar<uint8_t>},
    {{"short", "int16"}, 2, readScalar<int16_t>},
    {{"ushort", "uint16"}, 2, readScalar<uint16_t>},
    {{"int", "int32"}, 4, readScalar<int32_t>},
    {{"uint", "uint32"}, 4, readScalar<uint32_t>},
    {{"float", "float32"}, 4, readScalar<float>},
    {{"double", "float64"}, 8, readScalar<double>},
}};
constexpr int kFloat32 = 6;

int scalarIndex(std::string_view name) {
    for (size_t i = 0; i < kScalars.size(); ++i)
        if (name == kScalars[i].names[0] || name == kScalars[i].names[1]) return int(i);
    return -1;
}

// --- Stage 1: header -> schema -----------------------------------------------------------------

struct Failure {
    std::string text;
};

struct PropertySpec {
    std::string name;
    int scalar = kFloat32;  // index into kScalars; -1 for a list property
    uint32_t offset = 0;    // in the record (scalar properties)
};

struct ElementSpec {
    std::string name;
    uint64_t count = 0;
    uint32_t stride = 0;
    bool lists = false;  // has list properties (variable-size records)
    std::vector<PropertySpec> properties;
};

enum class Encoding { Unset, LittleEndian, BigEndian, Ascii };

struct Schema {
    Encoding encoding = Encoding::Unset;
    std::vector<ElementSpec> elements;
    std::vector<std::string> comments;
    uint64_t payloadOffset = 0;  // first byte after end_header
};

std::vector<std::string_view> splitWords(std::string_view s) {
    std::vector<std::string_view> w;
    size_t i = 0;
    while (true) {
        i = s.find_first_not_of(" \t", i);
        if (i == std::string_view::npos) break;
        const size_t j = std::min(s.find_first_of(" \t", i), s.size());
        w.push_back(s.substr(i, j - i));
        i = j;
    }
    return w;
}

bool toCount(std::string_view s, uint64_t& out) {
    if (s.empty() || s.size() > 19 || s.find_first_not_of("0123456789") != std::string_view::npos) return false;
    out = 0;
    for (char c : s) out = out * 10 + uint64_t(c - '0');
    return true;
}

constexpr size_t kHeaderLimit = size_t(1) << 20;  // 3DGS headers are ~1.5 KB

Schema readHeader(std::string_view bytes) {
    Schema schema;
    bool done = false;
    size_t lineNo = 0;
    std::string_view line;
    auto where = [&] { return "PLY header line " + std::to_string(lineNo) + ": "; };

    using Handler = std::function<void(const std::vector<std::string_view>&)>;
    const std::pair<std::string_view, Handler> handlers[] = {
        {"format",
         [&](const std::vector<std::string_view>& w) {
             if (w.size() < 2) throw Failure{where() + "malformed 'format' line"};
             static const std::pair<std::string_view, Encoding> encodings[] = {
                 {"binary_little_endian", Encoding::LittleEndian},
                 {"binary_big_endian", Encoding::BigEndian},
                 {"ascii", Encoding::Ascii}};
             for (const auto& [name, e] : encodings)
                 if (w[1] == name) {
                     schema.encoding = e;
                     return;
                 }
             throw Failure{where() + "unknown format '" + std::string(w[1]) + "'"};
         }},
        {"comment", [&](const std::vector<std::string_view>& w) {
             const size_t at = line.find(w[0]) + w[0].size();
             schema.comments.emplace_back(at < line.size() ? line.substr(at + 1) : std::string_view());
         }},
        {"obj_info", [&](const std::vector<std::string_view>& w) {
             const size_t at = line.find(w[0]) + w[0].size();
             schema.comments.emplace_back(at < line.size() ? line.substr(at + 1) : std::string_view());
         }},
        {"element",
         [&](const std::vector<std::string_view>& w) {
             if (w.size() != 3) throw Failure{where() + "malformed 'element' line"};
             ElementSpec e;
             e.name = std::string(w[1]);
             if (!toCount(w[2], e.count)) throw Failure{where() + "bad element count '" + std::string(w[2]) + "'"};
             schema.elements.push_back(std::move(e));
         }},
        {"property",
         [&](const std::vector<std::string_view>& w) {
             if (schema.elements.empty()) throw Failure{where() + "'property' before any 'element'"};
             ElementSpec& e = schema.elements.back();
             PropertySpec p;
             if (w.size() == 5 && w[1] == "list") {
                 if (scalarIndex(w[2]) < 0 || scalarIndex(w[3]) < 0) throw Failure{where() + "unknown list types"};
                 p.scalar = -1;
                 e.lists = true;
             } else if (w.size() == 3) {
                 p.scalar = scalarIndex(w[1]);
                 if (p.scalar < 0) throw Failure{where() + "unknown property type '" + std::string(w[1]) + "'"};
                 p.offset = e.stride;
                 e.stride += kScalars[size_t(p.scalar)].bytes;
             } else {
                 throw Failure{where() + "malformed 'property' line"};
             }
             p.name = std::string(w.back());
             e.properties.push_back(std::move(p));
         }},
        {"end_header",
         [&](const std::vector<std::string_view>&) {
             if (schema.encoding == Encoding::Unset) throw Failure{"PLY header has no 'format' line"};
             done = true;
         }},
    };

    const std::string_view head = bytes.substr(0, std::min(bytes.size(), kHeaderLimit));
    size_t cursor = 0;
    while (!done) {
        const size_t eol = head.find('\n', cursor);
        if (eol == std::string_view::npos)
            throw Failure{lineNo == 0 ? "not a PLY file (empty or no header)"
                                      : "PLY header has no 'end_header' line (file truncated or not a PLY)"};
        line = head.substr(cursor, eol - cursor);
        if (!line.empty() && line.back() == '\r') line.remove_suffix(1);  // CRLF headers
        cursor = eol + 1;
        ++lineNo;
        if (lineNo == 1) {
            if (line != "ply") throw Failure{"not a PLY file (first line is not 'ply')"};
            continue;
        }
        const std::vector<std::string_view> w = splitWords(line);
        if (w.empty()) continue;
        bool known = false;
        for (const auto& [keyword, handler] : handlers)
            if (w[0] == keyword) {
                handler(w);
                known = true;
                break;
            }
        if (!known) throw Failure{where() + "unknown keyword '" + std::string(w[0]) + "'"};
    }
    schema.payloadOffset = cursor;
    return schema;
}

// --- Stage 2: schema -> column plan --------------------------------------------------------------

// Gaussian attributes, by slot.
enum Attr : int { kX = 0, kDc = 3, kOpacity = 6, kScale = 7, kRot = 10, kRest = 14, kAttrCount = kRest + 45 };

// The fixed property names and their slots (f_rest_<j> is slot kRest + j).
constexpr std::pair<const char*, int> kNamedAttrs[] = {
    {"x", kX},          {"y", kX + 1},      {"z", kX + 2},      {"f_dc_0", kDc},     {"f_dc_1", kDc + 1},
    {"f_dc_2", kDc + 2}, {"opacity", kOpacity}, {"scale_0", kScale}, {"scale_1", kScale + 1},
    {"scale_2", kScale + 2}, {"rot_0", kRot}, {"rot_1", kRot + 1}, {"rot_2", kRot + 2}, {"rot_3", kRot + 3}};

struct Column {
    uint32_t offset;  // in the record
    Reader read;
    float* dst;       // value of gaussian i goes to dst[i * step]
    uint32_t step;
};

struct ColumnPlan {
    uint64_t count = 0, offset = 0;  // gaussians; file offset of the first record
    uint32_t stride = 0;
    int degree = 0;
    bool converted = false;  // some used property is not float32
    std::vector<Column> columns;
};

struct Output {
    gsr_ply_scene pub{};
    std::vector<float> posAlpha, scales, rotations, sh, rawOpacity, rawLogScales, rawRotations;
    std::vector<std::string> warningText, commentText;
    std::vector<const char*> warningPtrs, commentPtrs;
};

bool isCompressedFormat(const ElementSpec& e) {
    if (e.name == "chunk") return true;
    return std::any_of(e.properties.begin(), e.properties.end(),
                       [](const PropertySpec& p) { return p.name.compare(0, 7, "packed_") == 0; });
}

// Checks the schema, sizes the outputs and builds the column plan.
ColumnPlan planColumns(const Schema& schema, const std::string& file, uint64_t fileSize, Output& out) {
    switch (schema.encoding) {
        case Encoding::Ascii:
            throw Failure{file + ": ASCII PLY is not supported (3DGS files are binary_little_endian; re-export as binary)"};
        case Encoding::BigEndian:
            throw Failure{file + ": binary_big_endian PLY is not supported (expected binary_little_endian)"};
        default: break;
    }
    if (std::any_of(schema.elements.begin(), schema.elements.end(), isCompressedFormat))
        throw Failure{file + ": this is a PlayCanvas compressed PLY (element 'chunk' / packed_* properties), not a "
                             "standard 3DGS PLY; decompress it first"};
    const auto vertexIt = std::find_if(schema.elements.begin(), schema.elements.end(),
                                       [](const ElementSpec& e) { return e.name == "vertex"; });
    if (vertexIt == schema.elements.end()) throw Failure{file + ": no 'vertex' element"};
    const ElementSpec& vertex = *vertexIt;

    // Where the vertex records start, what follows them, and the elements that are skipped.
    ColumnPlan plan;
    plan.offset = schema.payloadOffset;
    uint64_t trailing = 0;
    bool trailingSized = true;
    for (auto it = schema.elements.begin(); it != schema.elements.end(); ++it) {
        if (it == vertexIt) continue;
        out.warningText.push_back("skipped element '" + it->name + "' (" + std::to_string(it->count) + " records)");
        if (it > vertexIt) {
            trailingSized = trailingSized && !it->lists;
            trailing += it->count * it->stride;
            continue;
        }
        if (it->lists) throw Failure{file + ": element '" + it->name + "' with list properties precedes 'vertex'"};
        const uint64_t room = fileSize > plan.offset ? fileSize - plan.offset : 0;
        if (it->stride && it->count > room / it->stride)
            throw Failure{file + ": file is truncated inside element '" + it->name + "'"};
        plan.offset += it->count * it->stride;
    }
    if (vertex.lists) throw Failure{file + ": 'vertex' element has list properties (not a 3DGS file)"};
    if (vertex.count == 0) throw Failure{file + ": 'vertex' element is empty (0 gaussians)"};
    if (vertex.count > 0xFFFFFFFFull) throw Failure{file + ": too many gaussians"};
    plan.count = vertex.count;
    plan.stride = vertex.stride;

    // Resolve every property to its slot (-1: not used by the scene).
    std::array<const PropertySpec*, kAttrCount> bySlot{};
    int restCount = 0;
    for (const PropertySpec& p : vertex.properties) {
        int slot = -1;
        for (const auto& [name, s] : kNamedAttrs)
            if (p.name == name) slot = s;
        if (slot < 0 && p.name.compare(0, 7, "f_rest_") == 0) {
            uint64_t j = 0;
            if (!toCount(std::string_view(p.name).substr(7), j) || j >= 45)
                throw Failure{file + ": bad SH property name '" + p.name + "'"};
            slot = kRest + int(j);
            ++restCount;
        }
        if (slot < 0) continue;
        if (bySlot[size_t(slot)]) throw Failure{file + ": duplicate property '" + p.name + "'"};
        bySlot[size_t(slot)] = &p;
    }
    for (const auto& [name, s] : kNamedAttrs)
        if (!bySlot[size_t(s)]) throw Failure{file + ": missing required property '" + std::string(name) + "'"};
    int highest = -1;
    for (int j = 0; j < 45; ++j)
        if (bySlot[size_t(kRest + j)]) highest = j;
    if (highest + 1 != restCount)
        throw Failure{file + ": f_rest_* indices are not contiguous (highest index " + std::to_string(highest) +
                      ", but only " + std::to_string(restCount) + " present)"};
    constexpr int kRestForDegree[4] = {0, 9, 24, 45};
    plan.degree = int(std::find(std::begin(kRestForDegree), std::end(kRestForDegree), restCount) -
                      std::begin(kRestForDegree));
    if (plan.degree > 3)
        throw Failure{file + ": unexpected number of f_rest_* properties: " + std::to_string(restCount) +
                      " (expected 0, 9, 24 or 45 for SH degree 0..3)"};

    // The payload must hold every record the header promises.
    const uint64_t room = fileSize > plan.offset ? fileSize - plan.offset : 0;
    const uint64_t need = plan.count * plan.stride;
    if (room < need)
        throw Failure{file + ": file is truncated: header declares " + std::to_string(plan.count) + " gaussians x " +
                      std::to_string(plan.stride) + " bytes = " + std::to_string(need) +
                      " bytes of vertex data, but only " + std::to_string(room) + " bytes follow the header (" +
                      std::to_string(plan.stride ? room / plan.stride : 0) + " complete records)"};
    if (trailingSized && room > need + trailing)
        out.warningText.push_back(std::to_string(room - need - trailing) +
                                  " unexpected trailing bytes after the last element");
    if (trailingSized && room < need + trailing) out.warningText.push_back("elements after 'vertex' are truncated (ignored)");

    // Outputs, then one column per used property with its destination.
    const size_t n = size_t(plan.count);
    const uint32_t K = uint32_t((plan.degree + 1) * (plan.degree + 1));
    out.posAlpha.resize(4 * n);
    out.scales.resize(4 * n);
    out.rotations.resize(4 * n);
    out.sh.resize(n * K * 3);
    out.rawOpacity.resize(n);
    out.rawLogScales.resize(3 * n);
    out.rawRotations.resize(4 * n);
    auto destination = [&](int slot, float*& dst, uint32_t& step) {
        if (slot < kDc) dst = &out.posAlpha[size_t(slot - kX)], step = 4;
        else if (slot < kOpacity) dst = &out.sh[size_t(slot - kDc)], step = K * 3;
        else if (slot == kOpacity) dst = out.rawOpacity.data(), step = 1;
        else if (slot < kRot) dst = &out.rawLogScales[size_t(slot - kScale)], step = 3;
        else if (slot < kRest) dst = &out.rawRotations[size_t(slot - kRot)], step = 4;
        else {
            // The file's f_rest is channel-major (j = c (K - 1) + k - 1); the scene's SH is
            // coefficient-major RGB triplets: coefficient k, channel c at k * 3 + c.
            const uint32_t j = uint32_t(slot - kRest), c = j / (K - 1), k = 1 + j % (K - 1);
            dst = &out.sh[size_t(k) * 3 + c];
            step = K * 3;
        }
    };
    for (int slot = 0; slot < kAttrCount; ++slot) {
        const PropertySpec* p = bySlot[size_t(slot)];
        if (!p) continue;
        Column c{p->offset, kScalars[size_t(p->scalar)].read, nullptr, 0};
        destination(slot, c.dst, c.step);
        plan.converted = plan.converted || p->scalar != kFloat32;
        plan.columns.push_back(c);
    }
    return plan;
}

// --- Stage 3: decode ---------------------------------------------------------------------------

bool notFinite(float v) { return !std::isfinite(v); }

constexpr float kMinQuatNorm2 = 1e-24f;  // at or below: the identity rotation
constexpr size_t kChunk = 8192;          // gaussians per work item

struct DecodeCounts {
    uint64_t nonFinite = 0;
    uint32_t zeroQuaternions = 0;
};

// Columns of gaussians [begin, end), then their activations.
void decodeChunk(const ColumnPlan& plan, const uint8_t* records, size_t begin, size_t end, Output& out,
                 DecodeCounts& counts) {
    for (const Column& c : plan.columns) {
        const uint8_t* src = records + begin * plan.stride + c.offset;
        float* dst = c.dst + begin * c.step;
        if (c.read == kScalars[kFloat32].read) {  // the common case: float32 copied as is
            for (size_t i = begin; i < end; ++i, src += plan.stride, dst += c.step) {
                std::memcpy(dst, src, 4);
                counts.nonFinite += notFinite(*dst) ? 1u : 0u;
            }
        } else {
            for (size_t i = begin; i < end; ++i, src += plan.stride, dst += c.step) {
                *dst = c.read(src);
                counts.nonFinite += notFinite(*dst) ? 1u : 0u;
            }
        }
    }
    for (size_t i = begin; i < end; ++i) {
        out.posAlpha[4 * i + 3] = 1.0f / (1.0f + std::exp(-out.rawOpacity[i]));
        for (int a = 0; a < 3; ++a) out.scales[4 * i + size_t(a)] = std::exp(out.rawLogScales[3 * i + size_t(a)]);
        out.scales[4 * i + 3] = 0.0f;
        const float* q = &out.rawRotations[4 * i];
        float* r = &out.rotations[4 * i];
        const float len2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        if (len2 > kMinQuatNorm2 && std::isfinite(len2)) {
            const float len = std::sqrt(len2);
            for (int a = 0; a < 4; ++a) r[a] = q[a] / len;
        } else {
            r[0] = 1.0f;  // w first
            r[1] = r[2] = r[3] = 0.0f;
            ++counts.zeroQuaternions;
        }
    }
}

gsr_status load(const uint8_t* bytes, uint64_t size, const std::string& file, gsr_ply_scene** result) {
    Output* out = new (std::nothrow) Output();
    if (!out) return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
    try {
        Schema schema;
        try {
            schema = readHeader(std::string_view(reinterpret_cast<const char*>(bytes), size_t(size)));
        } catch (const Failure& f) {
            throw Failure{file + ": " + f.text};
        }
        out->commentText = schema.comments;
        const ColumnPlan plan = planColumns(schema, file, size, *out);
        const size_t n = size_t(plan.count);
        const uint8_t* records = bytes + plan.offset;
        const size_t chunks = (n + kChunk - 1) / kChunk;
        const size_t workers =
            std::min<size_t>(chunks, std::max<size_t>(1, std::min<unsigned>(std::thread::hardware_concurrency(), 64)));
        std::vector<DecodeCounts> counts(workers);
        std::atomic<size_t> next{0};
        auto work = [&](size_t w) {
            for (size_t c = next++; c < chunks; c = next++)
                decodeChunk(plan, records, c * kChunk, std::min(n, (c + 1) * kChunk), *out, counts[w]);
        };
        std::vector<std::thread> pool;
        for (size_t w = 1; w < workers; ++w) pool.emplace_back(work, w);
        work(0);
        for (std::thread& t : pool) t.join();

        gsr_ply_scene& p = out->pub;
        for (const DecodeCounts& c : counts) {
            p.non_finite_values += c.nonFinite;
            p.zero_quaternions += c.zeroQuaternions;
        }
        if (p.non_finite_values)
            out->warningText.push_back(std::to_string(p.non_finite_values) + " NaN/Inf values in the file");
        if (p.zero_quaternions)
            out->warningText.push_back(std::to_string(p.zero_quaternions) +
                                       " zero-length quaternions replaced by identity");
        if (plan.converted) out->warningText.push_back("non-float32 properties converted to float");
        for (const std::string& w : out->warningText) out->warningPtrs.push_back(w.c_str());
        for (const std::string& c : out->commentText) out->commentPtrs.push_back(c.c_str());
        p.count = uint32_t(n);
        p.sh_degree = plan.degree;
        p.pos_alpha = out->posAlpha.data();
        p.scales = out->scales.data();
        p.rotations = out->rotations.data();
        p.sh = out->sh.data();
        p.raw_opacity = out->rawOpacity.data();
        p.raw_log_scales = out->rawLogScales.data();
        p.raw_rotations = out->rawRotations.data();
        p.warning_count = uint32_t(out->warningPtrs.size());
        p.warnings = out->warningPtrs.data();
        p.comment_count = uint32_t(out->commentPtrs.size());
        p.comments = out->commentPtrs.data();
        p.internal = out;
    } catch (const Failure& f) {
        delete out;
        return setError(GSR_ERROR_FORMAT, f.text.c_str());
    } catch (const std::bad_alloc&) {
        delete out;
        return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
    }
    *result = &out->pub;
    return GSR_OK;
}

std::string fileName(const char* path) {
    const std::string_view p(path);
    const size_t cut = p.find_last_of("/\\");
    return std::string(cut == std::string_view::npos ? p : p.substr(cut + 1));
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
    std::FILE* f = std::fopen(path, "rb");
    if (!f) return setError(GSR_ERROR_IO, (std::string("gsr_ply_load: cannot open ") + path).c_str());
    std::vector<uint8_t> bytes;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long size = ok ? std::ftell(f) : -1;
    ok = ok && size >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        try {
            bytes.resize(size_t(size));
        } catch (const std::bad_alloc&) {
            std::fclose(f);
            return setError(GSR_ERROR_OUT_OF_MEMORY, "gsr_ply_load: host allocation");
        }
        ok = bytes.empty() || std::fread(bytes.data(), 1, bytes.size(), f) == bytes.size();
    }
    std::fclose(f);
    if (!ok) return setError(GSR_ERROR_IO, (std::string("gsr_ply_load: cannot read ") + path).c_str());
    return load(bytes.data(), bytes.size(), fileName(path), out);
}

void gsr_ply_free(gsr_ply_scene* scene) {
    if (scene) delete static_cast<Output*>(scene->internal);
}

}  // extern "C"
