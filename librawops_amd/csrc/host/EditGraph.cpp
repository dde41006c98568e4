// Edit manifests, executable edit graphs and the device tile cache.
//
// The manifest format (JSON text, key order, number spelling, escapes, field
// names and validation rules) is the reference's (EditGraph.cpp:175-1536) so
// that saved edits and cache signatures carry over unchanged; the executable
// graph builds this library's nodes (EditGraph.cpp:1838-2215 restricted to
// linear chains of the render path's operations) and the cache keeps its
// tiles in HBM (EditGraph.cpp:23-173).
#include "../../../include/rawengine/EditGraph.hpp"
#include "../../../include/rawengine/ImageAnalysis.hpp"
#include "../../../include/rawengine/SpatialOps.hpp"
#include "../../../include/rawengine/ToneOps.hpp"
#include "internal.hpp"
#include "sha256.hpp"

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>
#include <set>
#include <stdexcept>
#include <type_traits>

namespace rawengine {

// ============================================================== tile cache ==

namespace {
constexpr std::size_t kEntryCharge = 256;  // bookkeeping allowance per entry (EditGraph.cpp:50)

void validate_cache_level(RenderLevel level) {
    if (level.mip > 31 || (level.quality != RenderQuality::Preview && level.quality != RenderQuality::Final))
        throw std::invalid_argument("invalid cache render level");
}
}  // namespace

TileCache::TileCache(std::size_t max_bytes) : max_bytes_(max_bytes) {}
TileCache::~TileCache() = default;

TileCache::Stats TileCache::stats() const {
    std::lock_guard lock(mutex_);
    return {entries_.size(), used_bytes_, hits_, misses_};
}

void TileCache::clear() {
    std::lock_guard lock(mutex_);
    ++generation_;
    entries_.clear();
    recency_.clear();
    used_bytes_ = hits_ = misses_ = 0;
}

namespace detail {

std::shared_ptr<const DeviceTile> CacheAccess::lookup(TileCache& c, const Signature& sig, Rect b, RenderLevel level,
                                                      std::uint64_t* generation) {
    validate_cache_level(level);
    const TileCache::Key key{sig, {b.x, b.y, b.width, b.height}, level.mip, level.quality, 0};
    std::lock_guard lock(c.mutex_);
    if (auto it = c.entries_.find(key); it != c.entries_.end()) {
        ++c.hits_;
        c.recency_.splice(c.recency_.begin(), c.recency_, it->second.recency);
        return it->second.tile;
    }
    ++c.misses_;
    *generation = c.generation_;
    return nullptr;
}

bool CacheAccess::admits(const TileCache& c, size_t floats) {
    return floats && floats <= (c.max_bytes_ - std::min(c.max_bytes_, kEntryCharge)) / sizeof(float);
}

void CacheAccess::store(TileCache& c, const Signature& sig, Rect b, RenderLevel level,
                        std::shared_ptr<const DeviceTile> tile, std::uint64_t generation) {
    if (!tile || !admits(c, tile->floats())) return;
    const size_t charge = kEntryCharge + tile->floats() * sizeof(float);
    const TileCache::Key key{sig, {b.x, b.y, b.width, b.height}, level.mip, level.quality, 0};
    std::lock_guard lock(c.mutex_);
    if (generation != c.generation_ || c.entries_.count(key)) return;
    while (c.used_bytes_ > c.max_bytes_ - charge) {
        const auto oldest = std::prev(c.recency_.end());
        c.used_bytes_ -= c.entries_.at(*oldest).charged_bytes;
        c.entries_.erase(*oldest);
        c.recency_.erase(oldest);
    }
    c.recency_.push_front(key);
    try {
        c.entries_.emplace(key, TileCache::Entry{std::move(tile), c.recency_.begin(), charge});
    } catch (...) {
        c.recency_.pop_front();
        throw;
    }
    c.used_bytes_ += charge;
}

}  // namespace detail

// EditGraph.cpp:25-80: the host-tile form.  A node this library compiles
// renders straight into a device entry; any other node's host tile is checked
// and uploaded.
Tile TileCache::render(const Node& node, std::array<std::uint8_t, 32> signature, Rect bounds, RenderLevel level) {
    validate_cache_level(level);
    if (!node.supports_level(level)) throw std::invalid_argument("node does not support this cache render level");
    const auto to_host = [&](const detail::DeviceTile& t) {
        Tile out{bounds, std::vector<float>(t.floats()), t.descriptor};
        if (!out.rgb.empty()) {
            const hipStream_t stream = detail::thread_stream();
            detail::hip(hipMemcpyAsync(out.rgb.data(), t.ptr, out.rgb.size() * sizeof(float), hipMemcpyDeviceToHost,
                                       stream), "hipMemcpyAsync(cache hit)");
            detail::hip(hipStreamSynchronize(stream), "stream synchronize");
        }
        return out;
    };
    std::uint64_t generation = 0;
    if (auto hit = detail::CacheAccess::lookup(*this, signature, bounds, level, &generation)) return to_host(*hit);
    const uint64_t pixels = uint64_t(bounds.width) * bounds.height;
    if (pixels > std::numeric_limits<std::size_t>::max() / 3) throw std::domain_error("node returned an invalid cache tile");
    if (detail::compilable(node) && pixels) {
        const Node* src = &node;
        while (src->input_node()) src = src->input_node();
        const auto sb = src->source_bounds();
        auto tile = std::make_shared<detail::DeviceTile>(bounds.width, bounds.height, node.output_descriptor());
        detail::render_to_device(node, sb ? *sb : bounds,
                                 RenderRequest{bounds, std::max(bounds.width, bounds.height), level},
                                 static_cast<float*>(tile->ptr), detail::thread_stream());
        detail::CacheAccess::store(*this, signature, bounds, level, tile, generation);
        return to_host(*tile);
    }
    Tile t = node.render_level(bounds, level);
    if (t.bounds.x != bounds.x || t.bounds.y != bounds.y || t.bounds.width != bounds.width ||
        t.bounds.height != bounds.height || t.descriptor != node.output_descriptor() ||
        t.rgb.size() != size_t(pixels) * 3)
        throw std::domain_error("node returned an invalid cache tile");
    if (!detail::CacheAccess::admits(*this, t.rgb.size())) return t;
    auto tile = std::make_shared<detail::DeviceTile>(bounds.width, bounds.height, t.descriptor);
    const hipStream_t stream = detail::thread_stream();
    detail::hip(hipMemcpyAsync(tile->ptr, t.rgb.data(), t.rgb.size() * sizeof(float), hipMemcpyHostToDevice, stream),
                "hipMemcpyAsync(cache entry)");
    detail::hip(hipStreamSynchronize(stream), "stream synchronize");
    detail::CacheAccess::store(*this, signature, bounds, level, tile, generation);
    return t;
}

// ==================================================================== JSON ==

namespace {

constexpr std::size_t kMaxManifestBytes = 16u << 20;
constexpr int kMaxDepth = 64;

[[noreturn]] void bad_json() { throw std::invalid_argument("invalid edit JSON"); }

// Well-formed UTF-8 (no overlong forms, surrogates or code points past U+10FFFF).
bool utf8_ok(std::string_view s) {
    size_t i = 0;
    while (i < s.size()) {
        const unsigned char b = static_cast<unsigned char>(s[i]);
        if (b < 0x80) { ++i; continue; }
        int extra;
        uint32_t cp, low;
        if ((b >> 5) == 0x6) { extra = 1; cp = b & 0x1f; low = 0x80; }
        else if ((b >> 4) == 0xe) { extra = 2; cp = b & 0x0f; low = 0x800; }
        else if ((b >> 3) == 0x1e) { extra = 3; cp = b & 0x07; low = 0x10000; }
        else return false;
        if (i + extra >= s.size()) return false;
        for (int k = 1; k <= extra; ++k) {
            const unsigned char t = static_cast<unsigned char>(s[i + k]);
            if ((t >> 6) != 0x2) return false;
            cp = cp << 6 | (t & 0x3f);
        }
        if (cp < low || cp > 0x10ffff || (cp >= 0xd800 && cp < 0xe000)) return false;
        i += size_t(extra) + 1;
    }
    return true;
}

void put_utf8(std::string& out, uint32_t cp) {
    if (cp < 0x80) {
        out += char(cp);
    } else if (cp < 0x800) {
        out += char(0xc0 | cp >> 6);
        out += char(0x80 | (cp & 0x3f));
    } else if (cp < 0x10000) {
        out += char(0xe0 | cp >> 12);
        out += char(0x80 | (cp >> 6 & 0x3f));
        out += char(0x80 | (cp & 0x3f));
    } else {
        out += char(0xf0 | cp >> 18);
        out += char(0x80 | (cp >> 12 & 0x3f));
        out += char(0x80 | (cp >> 6 & 0x3f));
        out += char(0x80 | (cp & 0x3f));
    }
}

// Recursive-descent reader.  Whitespace is space, tab, CR and LF; object keys
// must be unique; integers without fraction or exponent are int64, other
// numbers finite doubles.
class Reader {
public:
    explicit Reader(std::string_view text) : t_(text) {
        if (text.size() > kMaxManifestBytes || !utf8_ok(text))
            throw std::invalid_argument("edit JSON exceeds limit or is not UTF-8");
    }
    EditValue document() {
        EditValue v = value(0);
        skip();
        if (p_ != t_.size()) bad_json();
        return v;
    }

private:
    void skip() {
        while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\t' || t_[p_] == '\n' || t_[p_] == '\r')) ++p_;
    }
    bool eat(char ch) {
        skip();
        if (p_ < t_.size() && t_[p_] == ch) { ++p_; return true; }
        return false;
    }
    uint32_t hex4() {
        if (t_.size() - p_ < 4) bad_json();
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char ch = t_[p_++];
            v <<= 4;
            if (ch >= '0' && ch <= '9') v |= uint32_t(ch - '0');
            else if (ch >= 'a' && ch <= 'f') v |= uint32_t(ch - 'a' + 10);
            else if (ch >= 'A' && ch <= 'F') v |= uint32_t(ch - 'A' + 10);
            else bad_json();
        }
        return v;
    }
    std::string str() {
        if (!eat('"')) bad_json();
        std::string out;
        for (;;) {
            if (p_ >= t_.size()) bad_json();
            const char ch = t_[p_++];
            if (ch == '"') return out;
            if (static_cast<unsigned char>(ch) < 0x20) bad_json();
            if (ch != '\\') { out += ch; continue; }
            if (p_ >= t_.size()) bad_json();
            const char e = t_[p_++];
            switch (e) {
            case '"': case '\\': case '/': out += e; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                uint32_t cp = hex4();
                if (cp >= 0xdc00 && cp < 0xe000) bad_json();
                if (cp >= 0xd800 && cp < 0xdc00) {
                    if (t_.size() - p_ < 2 || t_[p_] != '\\' || t_[p_ + 1] != 'u') bad_json();
                    p_ += 2;
                    const uint32_t lo = hex4();
                    if (lo < 0xdc00 || lo >= 0xe000) bad_json();
                    cp = 0x10000 + ((cp - 0xd800) << 10) + (lo - 0xdc00);
                }
                put_utf8(out, cp);
                break;
            }
            default: bad_json();
            }
        }
    }
    bool digit() const { return p_ < t_.size() && t_[p_] >= '0' && t_[p_] <= '9'; }
    EditValue number() {
        const size_t begin = p_;
        if (t_[p_] == '-') ++p_;
        if (p_ >= t_.size()) bad_json();
        if (t_[p_] == '0') {
            ++p_;
        } else {
            if (!digit()) bad_json();
            while (digit()) ++p_;
        }
        bool real = false;
        if (p_ < t_.size() && t_[p_] == '.') {
            real = true;
            ++p_;
            if (!digit()) bad_json();
            while (digit()) ++p_;
        }
        if (p_ < t_.size() && (t_[p_] == 'e' || t_[p_] == 'E')) {
            real = true;
            ++p_;
            if (p_ < t_.size() && (t_[p_] == '+' || t_[p_] == '-')) ++p_;
            if (!digit()) bad_json();
            while (digit()) ++p_;
        }
        const char* a = t_.data() + begin;
        const char* z = t_.data() + p_;
        if (!real) {
            std::int64_t i = 0;
            const auto r = std::from_chars(a, z, i);
            if (r.ec != std::errc{} || r.ptr != z) bad_json();
            return EditValue{i};
        }
        double d = 0;
        const auto r = std::from_chars(a, z, d);
        if (r.ec != std::errc{} || r.ptr != z || !std::isfinite(d)) bad_json();
        return EditValue{d};
    }
    EditValue value(int depth) {
        if (depth > kMaxDepth) bad_json();
        skip();
        if (p_ >= t_.size()) bad_json();
        if (eat('{')) {
            EditValue::Object o;
            if (eat('}')) return EditValue{std::move(o)};
            for (;;) {
                skip();
                if (p_ >= t_.size() || t_[p_] != '"') bad_json();
                std::string k = str();
                if (!eat(':')) bad_json();
                if (!o.emplace(std::move(k), value(depth + 1)).second) bad_json();
                if (eat('}')) return EditValue{std::move(o)};
                if (!eat(',')) bad_json();
            }
        }
        if (eat('[')) {
            EditValue::Array a;
            if (eat(']')) return EditValue{std::move(a)};
            for (;;) {
                a.push_back(value(depth + 1));
                if (eat(']')) return EditValue{std::move(a)};
                if (!eat(',')) bad_json();
            }
        }
        if (t_[p_] == '"') return EditValue{str()};
        if (t_.substr(p_, 4) == "true") { p_ += 4; return EditValue{true}; }
        if (t_.substr(p_, 5) == "false") { p_ += 5; return EditValue{false}; }
        if (t_.substr(p_, 4) == "null") { p_ += 4; return EditValue{nullptr}; }
        if (t_[p_] == '-' || digit()) return number();
        bad_json();
    }

    std::string_view t_;
    size_t p_ = 0;
};

// Canonical writer: no whitespace, keys in map order, '"' and '\' escaped
// with a backslash and control bytes as \u00XX, everything else verbatim;
// doubles in shortest-general form with 17 significant digits and a ".0"
// suffix when that form has neither a point nor an exponent, zero as "0.0".
void write_string(std::string& out, std::string_view v) {
    if (!utf8_ok(v)) throw std::invalid_argument("edit string is not UTF-8");
    static constexpr char hexdig[] = "0123456789abcdef";
    out += '"';
    for (const char ch : v) {
        const unsigned char u = static_cast<unsigned char>(ch);
        if (ch == '"' || ch == '\\') {
            out += '\\';
            out += ch;
        } else if (u < 0x20) {
            out += "\\u00";
            out += hexdig[u >> 4];
            out += hexdig[u & 15];
        } else {
            out += ch;
        }
    }
    out += '"';
}

void write_json(std::string& out, const EditValue& v, int depth = 0) {
    if (depth > kMaxDepth) throw std::invalid_argument("edit value exceeds depth limit");
    std::visit([&](const auto& x) {
        using T = std::decay_t<decltype(x)>;
        if constexpr (std::is_same_v<T, std::nullptr_t>) {
            out += "null";
        } else if constexpr (std::is_same_v<T, bool>) {
            out += x ? "true" : "false";
        } else if constexpr (std::is_same_v<T, std::int64_t>) {
            out += std::to_string(x);
        } else if constexpr (std::is_same_v<T, double>) {
            if (!std::isfinite(x)) throw std::invalid_argument("non-finite edit parameter");
            if (x == 0.0) { out += "0.0"; return; }
            char buf[64];
            const auto r = std::to_chars(buf, buf + sizeof buf, x, std::chars_format::general, 17);
            if (r.ec != std::errc{}) throw std::invalid_argument("invalid edit number");
            const std::string_view num(buf, size_t(r.ptr - buf));
            out += num;
            if (num.find_first_of(".eE") == std::string_view::npos) out += ".0";
        } else if constexpr (std::is_same_v<T, std::string>) {
            write_string(out, x);
        } else if constexpr (std::is_same_v<T, EditValue::Array>) {
            out += '[';
            for (size_t i = 0; i < x.size(); ++i) {
                if (i) out += ',';
                write_json(out, x[i], depth + 1);
            }
            out += ']';
        } else {
            out += '{';
            bool first = true;
            for (const auto& [k, child] : x) {
                if (!first) out += ',';
                first = false;
                write_string(out, k);
                out += ':';
                write_json(out, child, depth + 1);
            }
            out += '}';
        }
    }, v.data);
}

// ---- typed field access (errors are std::invalid_argument) ----

EditValue::Object as_object(const EditValue& v) {
    if (auto* o = std::get_if<EditValue::Object>(&v.data)) return *o;
    throw std::invalid_argument("edit JSON object expected");
}
EditValue::Array as_array(const EditValue& v) {
    if (auto* a = std::get_if<EditValue::Array>(&v.data)) return *a;
    throw std::invalid_argument("edit JSON array expected");
}
EditValue take(EditValue::Object& o, const char* key) {
    auto it = o.find(key);
    if (it == o.end()) throw std::invalid_argument(std::string("missing edit field: ") + key);
    EditValue v = std::move(it->second);
    o.erase(it);
    return v;
}
std::string as_string(const EditValue& v) {
    if (auto* s = std::get_if<std::string>(&v.data)) return *s;
    throw std::invalid_argument("edit string expected");
}
uint32_t as_positive_u32(const EditValue& v) {
    auto* i = std::get_if<std::int64_t>(&v.data);
    if (!i || *i <= 0 || *i > std::int64_t(std::numeric_limits<uint32_t>::max()))
        throw std::invalid_argument("positive edit version expected");
    return uint32_t(*i);
}
bool as_bool(const EditValue& v) {
    if (auto* b = std::get_if<bool>(&v.data)) return *b;
    throw std::invalid_argument("edit boolean expected");
}
double as_number(const EditValue& v) {
    if (auto* d = std::get_if<double>(&v.data)) return *d;
    if (auto* i = std::get_if<std::int64_t>(&v.data)) return double(*i);
    throw std::invalid_argument("edit number expected");
}
std::map<std::string, std::string> as_edges(const EditValue& v) {
    std::map<std::string, std::string> out;
    for (const auto& [name, target] : as_object(v)) out.emplace(name, as_string(target));
    return out;
}
EditValue::Object edges_object(const std::map<std::string, std::string>& edges) {
    EditValue::Object o;
    for (const auto& [name, target] : edges) o.emplace(name, EditValue{target});
    return o;
}

// ---- enumerations and their spellings (EditGraph.cpp:478-534) ----

constexpr std::pair<WorkingSpace, const char*> kSpaces[] = {
    {WorkingSpace::LinearProPhotoD50, "linear_prophoto_d50"}, {WorkingSpace::LinearRec2020D65, "linear_rec2020_d65"}};
constexpr std::pair<EditSourceKind, const char*> kKinds[] = {
    {EditSourceKind::DecodedBayerU16, "decoded_bayer_u16"},
    {EditSourceKind::SceneLinearRasterF32, "scene_linear_raster_f32"},
    {EditSourceKind::IccRasterU16, "icc_raster_u16"},
    {EditSourceKind::CoverageRasterF32, "coverage_raster_f32"},
    {EditSourceKind::PremultipliedRgbaRasterF32, "premultiplied_rgba_raster_f32"}};
constexpr std::pair<EditDomain, const char*> kDomains[] = {
    {EditDomain::CameraLinear, "camera_linear"},
    {EditDomain::SceneLinearProPhotoD50, "scene_linear_prophoto_d50"},
    {EditDomain::SceneLinearRec2020D65, "scene_linear_rec2020_d65"},
    {EditDomain::SceneLinearSrgb, "scene_linear_srgb"},
    {EditDomain::ToneMappedUnmanaged, "tone_mapped_unmanaged"},
    {EditDomain::DisplayLinearSrgb, "display_linear_srgb"},
    {EditDomain::DisplayEncodedSrgb, "display_encoded_srgb"},
    {EditDomain::DisplayEncodedIcc, "display_encoded_icc"},
    {EditDomain::UnmanagedBounded, "unmanaged_bounded"},
    {EditDomain::Coverage, "coverage"},
    {EditDomain::PremultipliedSceneLinearProPhotoD50, "premultiplied_scene_linear_prophoto_d50"},
    {EditDomain::PremultipliedSceneLinearRec2020D65, "premultiplied_scene_linear_rec2020_d65"}};

template <class E, size_t N>
std::string name_of(const std::pair<E, const char*> (&table)[N], E e, const char* what) {
    for (const auto& [v, n] : table)
        if (v == e) return n;
    throw std::invalid_argument(std::string("unknown ") + what);
}
template <class E, size_t N>
E parse_name(const std::pair<E, const char*> (&table)[N], const std::string& s, const char* what) {
    for (const auto& [v, n] : table)
        if (s == n) return v;
    throw std::invalid_argument(std::string("unknown ") + what);
}
std::string space_name(WorkingSpace w) { return name_of(kSpaces, w, "working space"); }
std::string kind_name(EditSourceKind k) { return name_of(kKinds, k, "source kind"); }
std::string domain_name(EditDomain d) { return name_of(kDomains, d, "edit domain"); }

std::string digest_hex(const std::array<uint8_t, 32>& d) {
    static constexpr char hexdig[] = "0123456789abcdef";
    std::string s;
    for (uint8_t b : d) { s += hexdig[b >> 4]; s += hexdig[b & 15]; }
    return s;
}
std::array<uint8_t, 32> parse_digest(const std::string& s) {
    if (s.size() != 64) throw std::invalid_argument("invalid SHA-256 digest");
    std::array<uint8_t, 32> d{};
    for (size_t i = 0; i < 64; ++i) {
        const char ch = s[i];
        uint8_t nib;
        if (ch >= '0' && ch <= '9') nib = uint8_t(ch - '0');
        else if (ch >= 'a' && ch <= 'f') nib = uint8_t(ch - 'a' + 10);
        else throw std::invalid_argument("SHA-256 digest must be lowercase hex");
        d[i / 2] |= uint8_t(nib << (i % 2 ? 0 : 4));
    }
    return d;
}
bool any_set(const std::array<uint8_t, 32>& d) {
    return std::any_of(d.begin(), d.end(), [](uint8_t b) { return b != 0; });
}
// Lowercase 8-4-4-4-12 hex.
bool uuid_ok(const std::string& id) {
    if (id.size() != 36) return false;
    for (size_t i = 0; i < 36; ++i) {
        const bool dash = i == 8 || i == 13 || i == 18 || i == 23;
        const char ch = id[i];
        if (dash ? ch != '-' : !((ch >= '0' && ch <= '9') || (ch >= 'a' && ch <= 'f'))) return false;
    }
    return true;
}
// 1..128 characters of [A-Za-z0-9_.-].
bool name_ok(const std::string& n) {
    if (n.empty() || n.size() > 128) return false;
    for (const char ch : n)
        if (!((ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || (ch >= '0' && ch <= '9') || ch == '_' ||
              ch == '.' || ch == '-'))
            return false;
    return true;
}

void validate_icc(const IccProfileIdentity& icc) {
    const bool intent_ok = icc.intent == "perceptual" || icc.intent == "relative_colorimetric" ||
                           icc.intent == "saturation" || icc.intent == "absolute_colorimetric";
    if (!any_set(icc.profile_sha256) || !intent_ok || !name_ok(icc.engine) || !name_ok(icc.engine_version))
        throw std::invalid_argument("invalid ICC edit identity");
}
EditValue::Object icc_fields(const IccProfileIdentity& icc) {
    return {{"black_point_compensation", EditValue{icc.black_point_compensation}},
            {"engine", EditValue{icc.engine}},
            {"engine_version", EditValue{icc.engine_version}},
            {"intent", EditValue{icc.intent}},
            {"profile_sha256", EditValue{digest_hex(icc.profile_sha256)}}};
}
IccProfileIdentity read_icc(const EditValue& v) {
    auto f = as_object(v);
    IccProfileIdentity icc;
    icc.profile_sha256 = parse_digest(as_string(take(f, "profile_sha256")));
    icc.intent = as_string(take(f, "intent"));
    icc.black_point_compensation = as_bool(take(f, "black_point_compensation"));
    icc.engine = as_string(take(f, "engine"));
    icc.engine_version = as_string(take(f, "engine_version"));
    if (!f.empty()) throw std::invalid_argument("unknown ICC policy field");
    validate_icc(icc);
    return icc;
}

// Formats 1/2 Bayer sources mean bilinear version 1 (EditGraph.cpp:602-606).
EditSource canonical(EditSource s) {
    if (s.kind == EditSourceKind::DecodedBayerU16 && !s.demosaic) s.demosaic = RawDemosaicIdentity{};
    return s;
}

EditValue::Object source_fields(const EditSource& s) {
    EditValue::Object o{{"content_sha256", EditValue{digest_hex(s.content_sha256)}},
                        {"id", EditValue{s.id}},
                        {"kind", EditValue{kind_name(s.kind)}}};
    if (s.working_space) o.emplace("working_space", EditValue{space_name(*s.working_space)});
    if (s.icc_input) o.emplace("icc_input", EditValue{icc_fields(*s.icc_input)});
    if (s.demosaic)
        o.emplace("demosaic", EditValue{EditValue::Object{
                                  {"algorithm", EditValue{s.demosaic->algorithm}},
                                  {"processing_version", EditValue{std::int64_t(s.demosaic->processing_version)}}}});
    return o;
}
EditSource read_source(const EditValue& v) {
    auto f = as_object(v);
    EditSource s;
    s.id = as_string(take(f, "id"));
    s.kind = parse_name(kKinds, as_string(take(f, "kind")), "source kind");
    s.content_sha256 = parse_digest(as_string(take(f, "content_sha256")));
    if (f.count("working_space"))
        s.working_space = parse_name(kSpaces, as_string(take(f, "working_space")), "working space");
    if (f.count("icc_input")) s.icc_input = read_icc(take(f, "icc_input"));
    if (f.count("demosaic")) {
        auto d = as_object(take(f, "demosaic"));
        RawDemosaicIdentity id;
        id.algorithm = as_string(take(d, "algorithm"));
        id.processing_version = as_positive_u32(take(d, "processing_version"));
        if (!d.empty()) throw std::invalid_argument("unknown RAW demosaic policy field");
        s.demosaic = id;
    }
    if (!f.empty()) throw std::invalid_argument("unknown edit source field");
    return s;
}

EditValue::Object operation_fields(const EditOperation& op) {
    EditValue::Object o = op.extra_fields;
    const std::pair<const char*, EditValue> known[] = {
        {"blend_mode", EditValue{op.blend_mode}},
        {"enabled", EditValue{op.enabled}},
        {"id", EditValue{op.id}},
        {"input_domain", EditValue{domain_name(op.input_domain)}},
        {"inputs", EditValue{edges_object(op.inputs)}},
        {"masks", EditValue{edges_object(op.masks)}},
        {"opacity", EditValue{op.opacity}},
        {"output_domain", EditValue{domain_name(op.output_domain)}},
        {"parameters", EditValue{op.parameters}},
        {"processing_version", EditValue{std::int64_t(op.processing_version)}},
        {"schema_version", EditValue{std::int64_t(op.schema_version)}},
        {"type", EditValue{op.type_id}}};
    for (const auto& [k, v] : known)
        if (!o.emplace(k, v).second) throw std::invalid_argument("operation extension collides with known field");
    return o;
}
EditOperation read_operation(const EditValue& v) {
    auto f = as_object(v);
    EditOperation op;
    op.id = as_string(take(f, "id"));
    op.type_id = as_string(take(f, "type"));
    op.schema_version = as_positive_u32(take(f, "schema_version"));
    op.processing_version = as_positive_u32(take(f, "processing_version"));
    op.enabled = as_bool(take(f, "enabled"));
    op.input_domain = parse_name(kDomains, as_string(take(f, "input_domain")), "edit domain");
    op.output_domain = parse_name(kDomains, as_string(take(f, "output_domain")), "edit domain");
    op.parameters = as_object(take(f, "parameters"));
    op.inputs = as_edges(take(f, "inputs"));
    op.masks = as_edges(take(f, "masks"));
    op.blend_mode = as_string(take(f, "blend_mode"));
    op.opacity = as_number(take(f, "opacity"));
    op.extra_fields = std::move(f);
    return op;
}

// ---- operation families (EditGraph.cpp:972-1093) ----

bool one_of(const std::string& t, std::initializer_list<const char*> names) {
    for (const char* n : names)
        if (t == n) return true;
    return false;
}
bool mask_type(const std::string& t) {
    return one_of(t, {"rawengine.mask_invert", "rawengine.mask_combine", "rawengine.masked_mix",
                      "rawengine.mask_brush", "rawengine.mask_linear_gradient", "rawengine.mask_radial_gradient",
                      "rawengine.mask_polygon", "rawengine.mask_luminance_range", "rawengine.mask_color_range",
                      "rawengine.mask_depth_range", "rawengine.mask_density", "rawengine.mask_feather_binomial",
                      "rawengine.mask_refine_working_y"});
}
bool rgba_type(const std::string& t) {
    return one_of(t, {"rawengine.rgba_premultiply", "rawengine.rgba_apply_coverage", "rawengine.rgba_source_over",
                      "rawengine.rgba_alpha", "rawengine.rgba_straight_rgb"});
}
bool rgba_domain(EditDomain d) {
    return d == EditDomain::PremultipliedSceneLinearProPhotoD50 || d == EditDomain::PremultipliedSceneLinearRec2020D65;
}
// The render path's operations: these execute here.
bool path_type(const std::string& t) {
    return one_of(t, {"rawengine.white_balance", "rawengine.exposure", "rawengine.camera_to_working",
                      "rawengine.working_space_convert", "rawengine.working_to_srgb", "rawengine.tone_curve",
                      "rawengine.guided_filter_working_y", "rawengine.curves_extended", "rawengine.output_clip",
                      "rawengine.srgb_encode", "rawengine.icc_display", "rawengine.legacy.fixed_chain"});
}

double param(const EditOperation& op, const char* name) {
    auto it = op.parameters.find(name);
    if (it == op.parameters.end()) throw std::invalid_argument("missing operation parameter");
    return as_number(it->second);
}

// EditGraph.cpp:860-877.
ExtendedCurvesSettings curves_of(const EditValue::Object& p) {
    if (p.size() != 5 || !p.count("master") || !p.count("red") || !p.count("green") || !p.count("blue") ||
        !p.count("interpolation"))
        throw std::invalid_argument("extended curves need master/red/green/blue knots and interpolation only");
    ExtendedCurvesSettings s;
    const char* names[4] = {"master", "red", "green", "blue"};
    for (int c = 0; c < 4; ++c) {
        const auto pts = as_array(p.at(names[c]));
        if (pts.size() < 2 || pts.size() > 256) throw std::invalid_argument("extended curve needs 2..256 knots");
        auto& knots = c == 0 ? s.master.knots : s.channels[c - 1].knots;
        knots.clear();
        for (const auto& pt : pts) {
            const auto xy = as_array(pt);
            if (xy.size() != 2) throw std::invalid_argument("extended curve knot needs x and y");
            knots.push_back({as_number(xy[0]), as_number(xy[1])});
        }
    }
    const auto how = as_string(p.at("interpolation"));
    if (how == "linear") s.interpolation = CurveInterpolation::Linear;
    else if (how == "shape_preserving_cubic") s.interpolation = CurveInterpolation::ShapePreservingCubic;
    else throw std::invalid_argument("unsupported extended curve interpolation");
    validate_extended_curves_settings(s);
    return s;
}

// EditGraph.cpp:710-718.
WorkingYGuidedFilterSettings guided_of(const EditValue::Object& p) {
    if (p.size() != 2 || !p.count("radius") || !p.count("epsilon"))
        throw std::invalid_argument("guided filter needs radius and epsilon only");
    const double r = as_number(p.at("radius"));
    if (!std::isfinite(r) || r < 0 || r > 8 || std::floor(r) != r)
        throw std::invalid_argument("guided filter radius must be integral 0..8");
    WorkingYGuidedFilterSettings s{uint32_t(r), as_number(p.at("epsilon"))};
    validate_working_y_guided_filter_settings(s);
    return s;
}

// Parameter rules of the render path's operations (EditGraph.cpp:1094-1305,
// those rows).  Other operation types are carried without parameter checks.
void validate_path_parameters(const EditOperation& op) {
    const std::string& t = op.type_id;
    if (t == "rawengine.curves_extended") {
        if (op.schema_version != 1 || op.processing_version != kCurrentEditProcessingVersion)
            throw std::invalid_argument(
                "extended curves/gamma levels require schema 1 and processing version 2");
        (void)curves_of(op.parameters);
    } else if (t == "rawengine.exposure") {
        if (op.schema_version != 1 || op.parameters.size() != 1 || std::abs(param(op, "stops")) > 32.0)
            throw std::invalid_argument("invalid exposure operation");
    } else if (t == "rawengine.white_balance") {
        if (op.schema_version != 1 || op.parameters.size() != 3)
            throw std::invalid_argument("invalid white-balance operation");
        for (const char* n : {"red_gain", "green_gain", "blue_gain"}) {
            const double g = param(op, n);
            if (g <= 0.0 || g > 65536.0) throw std::invalid_argument("invalid white-balance gain");
        }
    } else if (t == "rawengine.tone_curve") {
        if (op.schema_version != 1 || op.parameters.size() != 2 || param(op, "shoulder") <= 0.0 ||
            param(op, "gamma") <= 0.0)
            throw std::invalid_argument("invalid tone-curve operation");
    } else if (t == "rawengine.guided_filter_working_y") {
        if (op.schema_version != 1 || op.processing_version != kCurrentEditProcessingVersion)
            throw std::invalid_argument("guided filter requires schema 1 and processing version 2");
        (void)guided_of(op.parameters);
        const bool working = op.input_domain == EditDomain::SceneLinearProPhotoD50 ||
                             op.input_domain == EditDomain::SceneLinearRec2020D65;
        if (op.input_domain != op.output_domain || !working || op.inputs.size() != 1 || !op.inputs.count("image") ||
            !op.masks.empty() || op.blend_mode != "normal" || op.opacity != 1 || !op.extra_fields.empty())
            throw std::invalid_argument(
                "guided filter requires a plain unary scene-linear working RGB operation");
    } else if (t == "rawengine.camera_to_working") {
        if (op.schema_version != 1 || op.parameters.size() != 1 || !op.parameters.count("matrix"))
            throw std::invalid_argument("invalid camera-to-working operation");
        const auto m = as_array(op.parameters.at("matrix"));
        if (m.size() != 9) throw std::invalid_argument("camera matrix needs nine values");
        for (const auto& v : m) (void)as_number(v);
    } else if (one_of(t, {"rawengine.working_space_convert", "rawengine.working_to_srgb", "rawengine.output_clip",
                          "rawengine.srgb_encode", "rawengine.icc_display"})) {
        if (op.schema_version != 1 || !op.parameters.empty())
            throw std::invalid_argument("invalid parameterless edit operation");
    } else if (t == "rawengine.legacy.fixed_chain") {
        if (op.schema_version != 1 || op.parameters.size() != 1 || !op.parameters.count("recipe") ||
            !std::holds_alternative<EditValue::Object>(op.parameters.at("recipe").data))
            throw std::invalid_argument("invalid legacy recipe snapshot");
    }
}

std::string manifest_json(const EditValue& v) {
    std::string s;
    write_json(s, v);
    return s;
}

// Kahn's order over operation edges, ready set seeded in id order.
std::vector<std::string> topological(const std::vector<EditOperation>& ops,
                                     std::map<std::string, std::vector<std::string>>* dependents_out = nullptr) {
    std::map<std::string, size_t> indegree;
    std::map<std::string, std::vector<std::string>> dependents;
    for (const auto& op : ops) indegree.emplace(op.id, 0);
    for (const auto& op : ops)
        for (const auto* edges : {&op.inputs, &op.masks})
            for (const auto& [name, target] : *edges)
                if (indegree.count(target)) {
                    ++indegree.at(op.id);
                    dependents[target].push_back(op.id);
                }
    std::queue<std::string> ready;
    for (const auto& [id, n] : indegree)
        if (!n) ready.push(id);
    std::vector<std::string> order;
    while (!ready.empty()) {
        std::string id = std::move(ready.front());
        ready.pop();
        for (const auto& d : dependents[id])
            if (--indegree.at(d) == 0) ready.push(d);
        order.push_back(std::move(id));
    }
    if (dependents_out) *dependents_out = std::move(dependents);
    return order;
}

}  // namespace

// ================================================================ manifest ==

// EditGraph.cpp:1309-1403.
void validate_edit_manifest(const EditManifest& m) {
    if (m.format_version < 2 || m.format_version > 5 || m.processing_version == 0 ||
        (m.working_space != WorkingSpace::LinearProPhotoD50 && m.working_space != WorkingSpace::LinearRec2020D65) ||
        m.sources.empty() || m.sources.size() > 100000 || m.operations.size() > 100000)
        throw std::invalid_argument("unsupported or incomplete edit manifest");
    std::set<std::string> ids;
    for (const auto& s : m.sources) {
        if (!uuid_ok(s.id) || !ids.insert(s.id).second || !any_set(s.content_sha256))
            throw std::invalid_argument("invalid or duplicate edit source identity");
        (void)kind_name(s.kind);
        const bool raw = s.kind == EditSourceKind::DecodedBayerU16;
        const bool coverage = s.kind == EditSourceKind::CoverageRasterF32;
        if (coverage && (m.format_version < 4 || m.processing_version != kCurrentEditProcessingVersion))
            throw std::invalid_argument("coverage source requires format 4 and processing version 2");
        if (s.kind == EditSourceKind::PremultipliedRgbaRasterF32 && (m.format_version != 5 || m.processing_version != 2))
            throw std::invalid_argument("RGBA source requires format5/process2");
        if (s.demosaic && (!raw || m.format_version < 3))
            throw std::invalid_argument("explicit demosaic requires a format-3-or-newer Bayer source");
        if (raw && m.format_version >= 3 && !s.demosaic)
            throw std::invalid_argument("format-3 Bayer source requires demosaic policy");
        if (s.demosaic) validate_raw_demosaic(*s.demosaic);
        if (s.icc_input) validate_icc(*s.icc_input);
        if ((s.kind == EditSourceKind::IccRasterU16) != s.icc_input.has_value())
            throw std::invalid_argument("ICC raster source requires input profile identity");
        if ((raw || coverage) == s.working_space.has_value())
            throw std::invalid_argument("raster source needs a working space; Bayer source cannot have one");
        if (s.working_space) (void)space_name(*s.working_space);
    }
    std::map<std::string, const EditOperation*> ops;
    for (const auto& op : m.operations) {
        if (!uuid_ok(op.id) || !ids.insert(op.id).second || !name_ok(op.type_id) || !op.schema_version ||
            !op.processing_version || op.processing_version > m.processing_version || !name_ok(op.blend_mode) ||
            !std::isfinite(op.opacity) || op.opacity < 0.0 || op.opacity > 1.0)
            throw std::invalid_argument("invalid or duplicate edit operation");
        (void)domain_name(op.input_domain);
        (void)domain_name(op.output_domain);
        for (const auto& [name, target] : op.inputs)
            if (!name_ok(name) || !uuid_ok(target)) throw std::invalid_argument("invalid edit input edge");
        for (const auto& [name, target] : op.masks)
            if (!name_ok(name) || !uuid_ok(target)) throw std::invalid_argument("invalid edit mask edge");
        validate_path_parameters(op);
        if ((rgba_type(op.type_id) || rgba_domain(op.input_domain) || rgba_domain(op.output_domain)) &&
            m.format_version != 5)
            throw std::invalid_argument("RGBA operations/domains require format5");
        if (mask_type(op.type_id) && m.format_version < 4)
            throw std::invalid_argument("mask operations require format 4");
        ops.emplace(op.id, &op);
    }
    if (!ids.count(m.output_id)) throw std::invalid_argument("edit output does not exist");
    if (m.output_profile) validate_icc(*m.output_profile);
    const auto out = ops.find(m.output_id);
    const bool icc_output = out != ops.end() && out->second->output_domain == EditDomain::DisplayEncodedIcc;
    if (icc_output != m.output_profile.has_value())
        throw std::invalid_argument("ICC output domain and profile identity must agree");
    for (const auto& op : m.operations)
        for (const auto* edges : {&op.inputs, &op.masks})
            for (const auto& [name, target] : *edges)
                if (!ids.count(target)) throw std::invalid_argument("edit edge target does not exist");
    if (topological(m.operations).size() != m.operations.size())
        throw std::invalid_argument("edit graph contains a cycle");
    // Free-form values must serialise (finite numbers, UTF-8 strings, depth).
    std::string scratch;
    write_json(scratch, EditValue{m.extra_fields});
    for (const auto& op : m.operations) {
        write_json(scratch, EditValue{op.parameters});
        write_json(scratch, EditValue{op.extra_fields});
    }
}

// EditGraph.cpp:1405-1435: records sorted by id, root keys in map order.
std::string serialize_edit_manifest(const EditManifest& m) {
    validate_edit_manifest(m);
    auto sources = m.sources;
    auto operations = m.operations;
    std::sort(sources.begin(), sources.end(), [](const auto& a, const auto& b) { return a.id < b.id; });
    std::sort(operations.begin(), operations.end(), [](const auto& a, const auto& b) { return a.id < b.id; });
    EditValue::Array src, ops;
    for (const auto& s : sources) src.push_back(EditValue{source_fields(s)});
    for (const auto& op : operations) ops.push_back(EditValue{operation_fields(op)});
    EditValue::Object root = m.extra_fields;
    const std::pair<const char*, EditValue> known[] = {
        {"format_version", EditValue{std::int64_t(m.format_version)}},
        {"operations", EditValue{std::move(ops)}},
        {"output", EditValue{m.output_id}},
        {"processing_version", EditValue{std::int64_t(m.processing_version)}},
        {"sources", EditValue{std::move(src)}},
        {"working_space", EditValue{space_name(m.working_space)}}};
    for (const auto& [k, v] : known)
        if (!root.emplace(k, v).second) throw std::invalid_argument("manifest extension collides with known field");
    if (m.output_profile && !root.emplace("output_profile", EditValue{icc_fields(*m.output_profile)}).second)
        throw std::invalid_argument("manifest output profile collision");
    std::string text = manifest_json(EditValue{std::move(root)});
    if (text.size() > kMaxManifestBytes) throw std::length_error("serialized edit manifest exceeds limit");
    return text;
}

// EditGraph.cpp:1437-1461: format 1 reads as 2 with raster sources given the
// document's working space.
EditManifest parse_edit_manifest(std::string_view json) {
    auto root = as_object(Reader(json).document());
    EditManifest m;
    const uint32_t saved = as_positive_u32(take(root, "format_version"));
    if (saved > 5) throw std::invalid_argument("unsupported edit manifest format version");
    m.format_version = saved == 1 ? 2 : saved;
    m.processing_version = as_positive_u32(take(root, "processing_version"));
    m.working_space = parse_name(kSpaces, as_string(take(root, "working_space")), "working space");
    m.output_id = as_string(take(root, "output"));
    for (const auto& v : as_array(take(root, "sources"))) m.sources.push_back(read_source(v));
    if (saved == 1)
        for (auto& s : m.sources)
            if (s.kind != EditSourceKind::DecodedBayerU16 && !s.working_space) s.working_space = m.working_space;
    for (const auto& v : as_array(take(root, "operations"))) m.operations.push_back(read_operation(v));
    if (root.count("output_profile")) m.output_profile = read_icc(take(root, "output_profile"));
    m.extra_fields = std::move(root);
    validate_edit_manifest(m);
    return m;
}

// EditGraph.cpp:1463-1536.
EditManifest snapshot_legacy_recipe(EditSource source, GraphRecipe recipe, LegacyRecipeEra era,
                                    std::string operation_id, std::optional<IccProfileIdentity> output_profile) {
    EditManifest m;
    if (era == LegacyRecipeEra::ImplicitRec2020) {
        m.processing_version = kLegacyRec2020ProcessingVersion;
        m.working_space = WorkingSpace::LinearRec2020D65;
    } else if (era == LegacyRecipeEra::ImplicitProPhoto) {
        m.processing_version = kCurrentEditProcessingVersion;
        m.working_space = WorkingSpace::LinearProPhotoD50;
    } else {
        throw std::invalid_argument("unknown legacy recipe provenance");
    }
    const bool icc = recipe.output_mode == OutputMode::IccDisplay;
    if (icc && !output_profile) throw std::invalid_argument("legacy ICC recipe needs exact output profile identity");
    if (!icc && output_profile) throw std::invalid_argument("unexpected legacy output profile identity");
    if (recipe.camera_color && recipe.camera_color->target != m.working_space)
        throw std::invalid_argument("legacy camera target conflicts with declared provenance");
    const bool raw = source.kind == EditSourceKind::DecodedBayerU16;
    if (!raw && (recipe.red_gain != 1.0f || recipe.green_gain != 1.0f || recipe.blue_gain != 1.0f ||
                 recipe.camera_color))
        throw std::invalid_argument("legacy raster recipe has RAW calibration controls");
    if (source.working_space && *source.working_space != m.working_space)
        throw std::invalid_argument("legacy raster source space conflicts with declared provenance");
    if (raw && !recipe.camera_color && recipe.output_mode != OutputMode::LegacyBounded)
        throw std::invalid_argument("legacy RAW preview requires camera color calibration");
    m.sources.push_back(std::move(source));
    m.output_profile = std::move(output_profile);
    EditOperation op;
    op.id = std::move(operation_id);
    op.type_id = "rawengine.legacy.fixed_chain";
    op.processing_version = m.processing_version;
    op.input_domain = raw ? EditDomain::CameraLinear
                          : m.working_space == WorkingSpace::LinearProPhotoD50 ? EditDomain::SceneLinearProPhotoD50
                                                                               : EditDomain::SceneLinearRec2020D65;
    std::string mode;
    switch (recipe.output_mode) {
    case OutputMode::LegacyBounded: op.output_domain = EditDomain::UnmanagedBounded; mode = "legacy_bounded"; break;
    case OutputMode::SrgbPreview: op.output_domain = EditDomain::DisplayEncodedSrgb; mode = "srgb_preview"; break;
    case OutputMode::IccDisplay: op.output_domain = EditDomain::DisplayEncodedIcc; mode = "icc_display"; break;
    default: throw std::invalid_argument("unknown legacy output mode");
    }
    EditValue::Object fields{{"red_gain", EditValue{double(recipe.red_gain)}},
                             {"green_gain", EditValue{double(recipe.green_gain)}},
                             {"blue_gain", EditValue{double(recipe.blue_gain)}},
                             {"exposure_stops", EditValue{double(recipe.exposure_stops)}},
                             {"tone_shoulder", EditValue{double(recipe.tone_shoulder)}},
                             {"tone_gamma", EditValue{double(recipe.tone_gamma)}},
                             {"output_mode", EditValue{mode}}};
    if (recipe.camera_color) {
        EditValue::Array mat;
        for (double v : recipe.camera_color->camera_to_xyz_d50) mat.push_back(EditValue{v});
        fields.emplace("camera_to_xyz_d50", EditValue{std::move(mat)});
        fields.emplace("camera_target", EditValue{space_name(m.working_space)});
    }
    op.parameters.emplace("recipe", EditValue{std::move(fields)});
    op.inputs.emplace("image", m.sources.front().id);
    m.output_id = op.id;
    m.operations.push_back(std::move(op));
    validate_edit_manifest(m);
    return m;
}

// ============================================================= signatures ==

namespace detail {

namespace {
Signature signature_of(const EditValue::Object& identity, const Signature* upstream,
                       const std::optional<IccProfileIdentity>& profile) {
    static constexpr char kTag[] = "librawops.tile-cache.v2";  // hashed with its terminating NUL
    Sha256 h;
    h.update(kTag, sizeof kTag);
    if (upstream) h.update(upstream->data(), upstream->size());
    const std::string text = manifest_json(EditValue{identity});
    h.update(text.data(), text.size());
    if (profile) {
        const std::string p = manifest_json(EditValue{icc_fields(*profile)});
        h.update(p.data(), p.size());
    }
    return h.finish();
}
}  // namespace

Signature signature_of_operation(const EditOperation& op, const Signature& upstream,
                                 const std::optional<IccProfileIdentity>& output_profile) {
    return signature_of(operation_fields(op), &upstream, output_profile);
}

std::array<std::uint8_t, 32> edit_operation_signature(const EditOperation& op,
                                                      const std::array<std::uint8_t, 32>& upstream) {
    return signature_of_operation(op, upstream, std::nullopt);
}

std::array<std::uint8_t, 32> edit_source_signature(const EditSource& source) {
    return signature_of(source_fields(canonical(source)), nullptr, std::nullopt);
}

}  // namespace detail

// ========================================================= executable graph ==

namespace {

EditDomain domain_of(const ImageDescriptor& d) {
    if (d == ImageDescriptor::camera_linear()) return EditDomain::CameraLinear;
    if (d == ImageDescriptor::scene_linear(WorkingSpace::LinearProPhotoD50)) return EditDomain::SceneLinearProPhotoD50;
    if (d == ImageDescriptor::scene_linear(WorkingSpace::LinearRec2020D65)) return EditDomain::SceneLinearRec2020D65;
    if (d == ImageDescriptor::linear_srgb()) return EditDomain::SceneLinearSrgb;
    if (d.domain == PixelDomain::ToneMappedUnmanagedRGB && d.transfer == TransferFunction::CustomTone)
        return EditDomain::ToneMappedUnmanaged;
    if (d == ImageDescriptor::display_linear_srgb()) return EditDomain::DisplayLinearSrgb;
    if (d == ImageDescriptor::srgb_output()) return EditDomain::DisplayEncodedSrgb;
    if (d.domain == PixelDomain::DisplayEncodedRGB && d.primaries == ColorPrimaries::ICCProfile &&
        d.transfer == TransferFunction::ICCProfile)
        return EditDomain::DisplayEncodedIcc;
    if (d.domain == PixelDomain::BoundedUnmanagedRGB && d.transfer == TransferFunction::CustomTone)
        return EditDomain::UnmanagedBounded;
    throw std::invalid_argument("runtime node has an unsupported edit color domain");
}

WorkingSpace working_of(EditDomain d) {
    if (d == EditDomain::SceneLinearProPhotoD50) return WorkingSpace::LinearProPhotoD50;
    if (d == EditDomain::SceneLinearRec2020D65) return WorkingSpace::LinearRec2020D65;
    throw std::invalid_argument("operation target is not a supported working space");
}

float scalar(const EditValue::Object& p, const char* key) {
    auto it = p.find(key);
    if (it == p.end()) throw std::invalid_argument("missing operation scalar");
    const double v = as_number(it->second);
    if (!std::isfinite(v) || std::abs(v) > std::numeric_limits<float>::max())
        throw std::invalid_argument("operation scalar exceeds float32 range");
    return float(v);
}

CameraColorTransform camera_of(const EditValue::Object& p, WorkingSpace target, const char* key) {
    auto it = p.find(key);
    if (it == p.end()) throw std::invalid_argument("missing camera color matrix");
    const auto values = as_array(it->second);
    if (values.size() != 9) throw std::invalid_argument("camera matrix needs nine values");
    CameraColorTransform t;
    t.target = target;
    for (size_t i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = as_number(values[i]);
    return t;
}

// EditGraph.cpp:1613-1653.
std::shared_ptr<const Node> legacy_chain(const EditOperation& op, std::shared_ptr<const Node> source,
                                         WorkingSpace space, const std::shared_ptr<const IccDisplayTransform>& icc) {
    auto f = as_object(op.parameters.at("recipe"));
    const float red = scalar(f, "red_gain"), green = scalar(f, "green_gain"), blue = scalar(f, "blue_gain");
    const float stops = scalar(f, "exposure_stops");
    const float shoulder = scalar(f, "tone_shoulder"), gamma = scalar(f, "tone_gamma");
    const std::string mode = as_string(take(f, "output_mode"));
    std::shared_ptr<const Node> n = source;
    if (domain_of(source->output_descriptor()) == EditDomain::CameraLinear) {
        n = std::make_shared<WhiteBalanceNode>(n, red, green, blue);
        n = std::make_shared<ExposureNode>(n, stops);
        if (f.count("camera_to_xyz_d50")) {
            if (as_string(take(f, "camera_target")) != space_name(space))
                throw std::invalid_argument("legacy camera target disagrees with working space");
            n = std::make_shared<CameraToWorkingNode>(n, camera_of(f, space, "camera_to_xyz_d50"));
            f.erase("camera_to_xyz_d50");
        }
    } else {
        if (red != 1.0f || green != 1.0f || blue != 1.0f || f.count("camera_to_xyz_d50") || f.count("camera_target"))
            throw std::invalid_argument("legacy raster recipe contains RAW controls");
        n = std::make_shared<ExposureNode>(n, stops);
    }
    if (mode == "srgb_preview" || mode == "icc_display") n = std::make_shared<WorkingToSrgbNode>(n);
    else if (mode != "legacy_bounded") throw std::invalid_argument("unknown legacy output mode");
    if (f.size() != 6) throw std::invalid_argument("legacy recipe contains unrecognized fields");
    auto tone = std::make_shared<ToneCurveNode>(n, shoulder, gamma);
    if (mode == "srgb_preview") return std::make_shared<SrgbEncodeNode>(tone);
    if (mode == "icc_display") return std::make_shared<IccDisplayNode>(tone, icc);
    return std::make_shared<OutputClipNode>(tone);
}

// EditGraph.cpp:1655-1734, the render path's rows.
std::shared_ptr<const Node> build_node(const EditOperation& op, std::shared_ptr<const Node> in, WorkingSpace space,
                                       Rect bounds, const std::shared_ptr<const IccDisplayTransform>& icc) {
    const auto& p = op.parameters;
    const std::string& t = op.type_id;
    if (t == "rawengine.white_balance")
        return std::make_shared<WhiteBalanceNode>(in, scalar(p, "red_gain"), scalar(p, "green_gain"),
                                                  scalar(p, "blue_gain"));
    if (t == "rawengine.exposure") return std::make_shared<ExposureNode>(in, scalar(p, "stops"));
    if (t == "rawengine.camera_to_working")
        return std::make_shared<CameraToWorkingNode>(in, camera_of(p, working_of(op.output_domain), "matrix"), bounds);
    if (t == "rawengine.working_space_convert")
        return std::make_shared<WorkingSpaceConvertNode>(in, working_of(op.output_domain));
    if (t == "rawengine.working_to_srgb") return std::make_shared<WorkingToSrgbNode>(in);
    if (t == "rawengine.tone_curve")
        return std::make_shared<ToneCurveNode>(in, scalar(p, "shoulder"), scalar(p, "gamma"));
    if (t == "rawengine.guided_filter_working_y")
        return std::make_shared<WorkingYGuidedFilterNode>(in, bounds, guided_of(p));
    if (t == "rawengine.curves_extended") return std::make_shared<ExtendedCurvesNode>(in, curves_of(p));
    if (t == "rawengine.output_clip") return std::make_shared<OutputClipNode>(in);
    if (t == "rawengine.srgb_encode") return std::make_shared<SrgbEncodeNode>(in);
    if (t == "rawengine.icc_display") return std::make_shared<IccDisplayNode>(in, icc);
    if (t == "rawengine.legacy.fixed_chain") return legacy_chain(op, in, space, icc);
    throw std::invalid_argument("unknown operation cannot execute: " + t);
}

bool addressable(Rect b) {
    return b.width && b.height && uint64_t(b.x) + b.width <= (uint64_t{1} << 32) &&
           uint64_t(b.y) + b.height <= (uint64_t{1} << 32);
}

}  // namespace

// EditGraph.cpp:1838-2215 for linear chains over one RGB source each.
ExecutableEditGraph::ExecutableEditGraph(EditManifest manifest, std::vector<BoundEditSource> sources,
                                         std::shared_ptr<const IccDisplayTransform> icc,
                                         std::shared_ptr<TileCache> cache)
    : manifest_(std::move(manifest)) {
    validate_edit_manifest(manifest_);
    const auto& m = manifest_;
    if (m.processing_version != kLegacyRec2020ProcessingVersion && m.processing_version != kCurrentEditProcessingVersion)
        throw std::invalid_argument("unsupported edit processing version");
    if (m.processing_version == kLegacyRec2020ProcessingVersion && m.working_space != WorkingSpace::LinearRec2020D65)
        throw std::invalid_argument("legacy process version requires Rec.2020");
    if (icc) {
        if (!m.output_profile || icc->output_icc_identity() != m.output_profile ||
            icc->profile_sha256() != m.output_profile->profile_sha256)
            throw std::invalid_argument("runtime ICC output policy differs from manifest");
    } else if (m.output_profile) {
        throw std::invalid_argument("ICC output transform is missing");
    }

    struct Built {
        std::shared_ptr<const Node> node;
        Rect bounds;
        detail::Signature signature;
    };
    std::map<std::string, Built> built;
    if (sources.size() != m.sources.size()) throw std::invalid_argument("source binding count differs from manifest");
    std::map<std::string, const BoundEditSource*> bindings;
    for (const auto& b : sources)
        if (!bindings.emplace(b.identity.id, &b).second) throw std::invalid_argument("duplicate source binding");
    for (const auto& record : m.sources) {
        const auto found = bindings.find(record.id);
        if (found == bindings.end()) throw std::invalid_argument("source binding is missing");
        const BoundEditSource& b = *found->second;
        if (record.kind == EditSourceKind::CoverageRasterF32 || record.kind == EditSourceKind::PremultipliedRgbaRasterF32)
            throw std::invalid_argument("coverage and RGBA sources are outside the device render path");
        if (canonical(b.identity) != canonical(record) || !b.node || !addressable(b.bounds))
            throw std::invalid_argument("invalid source binding or bounds");
        if (b.node->source_fingerprint() != record.content_sha256)
            throw std::invalid_argument("runtime source fingerprint differs from manifest");
        const auto actual = b.node->source_bounds();
        if (!actual || actual->x != b.bounds.x || actual->y != b.bounds.y || actual->width != b.bounds.width ||
            actual->height != b.bounds.height)
            throw std::invalid_argument("runtime source bounds differ from binding");
        const EditDomain expected = record.kind == EditSourceKind::DecodedBayerU16 ? EditDomain::CameraLinear
                                    : *record.working_space == WorkingSpace::LinearProPhotoD50
                                        ? EditDomain::SceneLinearProPhotoD50
                                        : EditDomain::SceneLinearRec2020D65;
        if (domain_of(b.node->output_descriptor()) != expected || b.node->input_icc_identity() != record.icc_input ||
            b.node->raw_demosaic_identity() != canonical(record).demosaic)
            throw std::invalid_argument("runtime source color identity differs from manifest");
        const auto sig = detail::edit_source_signature(record);
        // A distinct node per record keeps the record's identity even when two
        // records share one runtime node.
        auto node = std::make_shared<detail::CachedNode>(b.node, cache, sig);
        source_nodes_.emplace(node.get(), std::make_pair(record.id, b.bounds));
        built.emplace(record.id, Built{std::move(node), b.bounds, sig});
    }

    std::map<std::string, const EditOperation*> ops;
    for (const auto& op : m.operations) ops.emplace(op.id, &op);
    for (const auto& id : topological(m.operations)) {
        const EditOperation& op = *ops.at(id);
        if (!path_type(op.type_id)) {
            if (mask_type(op.type_id) || rgba_type(op.type_id) ||
                one_of(op.type_id, {"rawengine.linear_mix", "rawengine.masked_mix"}) ||
                op.type_id.rfind("rawengine.", 0) == 0)
                throw std::invalid_argument("operation is outside the device render path: " + op.type_id);
            throw std::invalid_argument("unknown operation cannot execute: " + op.type_id);
        }
        if (op.processing_version != m.processing_version || op.schema_version != 1)
            throw std::invalid_argument("operation processing/schema version is unsupported");
        if (op.inputs.size() != 1 || !op.inputs.count("image") || !op.masks.empty() || op.blend_mode != "normal" ||
            op.opacity != 1.0 || !op.extra_fields.empty())
            throw std::invalid_argument("operation uses an unsupported input, mask, blend or extension");
        const Built& up = built.at(op.inputs.at("image"));
        if (domain_of(up.node->output_descriptor()) != op.input_domain)
            throw std::invalid_argument("operation input domain differs from runtime edge");
        std::shared_ptr<const Node> node;
        if (op.enabled) {
            node = build_node(op, up.node, m.working_space, up.bounds, icc);
        } else {
            if (op.output_domain != op.input_domain)
                throw std::invalid_argument("disabled operation cannot change color domain");
            node = up.node;
        }
        if (domain_of(node->output_descriptor()) != op.output_domain)
            throw std::invalid_argument("operation output domain differs from runtime node");
        auto sig = up.signature;
        if (op.enabled) {
            sig = detail::signature_of_operation(op, up.signature,
                                                 op.output_domain == EditDomain::DisplayEncodedIcc ? m.output_profile
                                                                                                   : std::nullopt);
            if (cache) node = std::make_shared<detail::CachedNode>(node, cache, sig);
        }
        built.emplace(id, Built{std::move(node), up.bounds, sig});
    }
    const auto out = built.find(m.output_id);
    if (out == built.end()) throw std::invalid_argument("edit output could not be constructed");
    output_ = out->second.node;
    output_bounds_ = out->second.bounds;
}

Rect ExecutableEditGraph::source_bounds() const {
    const Node* n = output_.get();
    while (n && !source_nodes_.count(n)) n = n->input_node();
    if (!n) throw std::invalid_argument("RGB node has no registered source or inputs");
    return source_nodes_.at(n).second;
}

Rect ExecutableEditGraph::required_source_region(Rect output) const { return required_source_region(output, {}); }

Rect ExecutableEditGraph::required_source_region(Rect output, RenderLevel level) const {
    const auto regions = required_source_regions(output, level);
    if (regions.size() != 1) throw std::invalid_argument("graph has multiple sources; use required_source_regions");
    return regions.begin()->second;
}

// EditGraph.cpp:2292-2429 for a linear chain: each node maps its output region
// back through input_region_level at its input level; the source publishes the
// native rectangle.
std::map<std::string, Rect> ExecutableEditGraph::required_source_regions(Rect output, RenderLevel level) const {
    auto level_bounds = [](Rect b, RenderLevel l) {
        if (l.mip == 0 && (l.quality == RenderQuality::Final || l.quality == RenderQuality::Preview)) return b;
        if (l.mip < 1 || l.mip > 2 || l.quality != RenderQuality::Preview)
            throw std::invalid_argument("unsupported planned render level");
        const uint32_t s = 1u << l.mip;
        return Rect{0, 0, b.width / s + (b.width % s != 0), b.height / s + (b.height % s != 0)};
    };
    if (!output_->supports_level(level)) throw std::invalid_argument("graph does not support this planned render level");
    const Rect extent = level_bounds(output_bounds_, level);
    if (output.x < extent.x || output.y < extent.y ||
        uint64_t(output.x) + output.width > uint64_t(extent.x) + extent.width ||
        uint64_t(output.y) + output.height > uint64_t(extent.y) + extent.height)
        throw std::out_of_range("planned ROI is outside graph output bounds");
    std::map<std::string, Rect> result;
    const Node* n = output_.get();
    Rect region = output;
    RenderLevel l = level;
    for (;;) {
        const RenderLevel up = n->input_level(l);
        const Rect required = n->input_region_level(region, level_bounds(output_bounds_, up), l);
        if (auto s = source_nodes_.find(n); s != source_nodes_.end()) {
            const Rect b = s->second.second;
            Rect native = required;
            if (up.mip > 0) {
                const uint32_t sc = 1u << up.mip;
                const auto left = std::min<uint64_t>(uint64_t(required.x) * sc, b.width);
                const auto top = std::min<uint64_t>(uint64_t(required.y) * sc, b.height);
                const auto right = std::min<uint64_t>((uint64_t(required.x) + required.width) * sc, b.width);
                const auto bottom = std::min<uint64_t>((uint64_t(required.y) + required.height) * sc, b.height);
                native = {uint32_t(b.x + left), uint32_t(b.y + top), uint32_t(right - left), uint32_t(bottom - top)};
            }
            result.emplace(s->second.first, native);
            return result;
        }
        const Node* next = n->input_node();
        if (!next) throw std::invalid_argument("runtime node has no registered source or input mapping");
        n = next;
        region = required;
        l = up;
    }
}

// ================================================================ renderer ==

void Renderer::render_tiles(const ExecutableEditGraph& graph, RenderRequest request, const TileCallback& callback,
                            const CancellationToken* cancellation) const {
    render_tiles(graph.output(), graph.output_bounds(), request, callback, cancellation);
}

void Renderer::render_tiles(const ExecutableEditGraph& graph, Rect viewport, const TileCallback& callback,
                            std::uint32_t tile_size, const CancellationToken* cancellation) const {
    render_tiles(graph.output(), graph.output_bounds(), viewport, callback, tile_size, cancellation);
}

Tile Renderer::render_image(const ExecutableEditGraph& graph, RenderRequest request,
                            const CancellationToken* cancellation) const {
    return render_image(graph.output(), graph.output_bounds(), request, cancellation);
}

Tile Renderer::render_image(const ExecutableEditGraph& graph, Rect viewport, std::uint32_t tile_size,
                            const CancellationToken* cancellation) const {
    return render_image(graph.output(), graph.output_bounds(), viewport, tile_size, cancellation);
}

std::vector<float> Renderer::render_roi(const ExecutableEditGraph& graph, Rect viewport, std::uint32_t tile_size,
                                        const CancellationToken* cancellation) const {
    return std::move(render_image(graph, viewport, tile_size, cancellation).rgb);
}

RgbHistogram histogram_rgb(const ExecutableEditGraph& graph, const RenderRequest& request,
                           const HistogramOptions& options, const CancellationToken* cancellation) {
    return histogram_rgb(graph.output(), graph.output_bounds(), request, options, cancellation);
}

}  // namespace rawengine
