// Saved edit manifests, their executable graphs and the tile cache, MI355X
// build.  Source-compatible subset of the reference's EditGraph.hpp
// (tylermcm/librawops EditGraph.hpp:18-210) for the tiled render path:
//
//   * EditValue / EditSource / EditOperation / EditManifest with the same JSON
//     format (deterministic key order, format versions 1..5 readable, unknown
//     operations and extension fields carried through parse/save);
//   * ExecutableEditGraph for linear chains of the render path's operations
//     (white_balance, exposure, camera_to_working, working_space_convert,
//     working_to_srgb, tone_curve, guided_filter_working_y, curves_extended,
//     output_clip, srgb_encode, icc_display, legacy.fixed_chain) over a
//     decoded Bayer or scene-linear raster source.  The chain is compiled into
//     device plans: point operations fuse into the demosaic kernel's op
//     program, a guided filter starts a second launch over the first one's
//     device output.  Other operation types parse and serialise but do not
//     execute here (std::invalid_argument at graph construction);
//   * TileCache keyed by the reference's SHA-256 signatures, holding tiles in
//     HBM.  Entries are made where a compiled chain already ends a launch (the
//     input of a guided filter and a chain's output), not at every node: inside
//     one fused launch an upstream tile costs more HBM bytes to read back than
//     to recompute from the 2-byte Bayer samples.
//
// Not here (outside the render path): coverage and premultiplied-RGBA payloads,
// masks, multi-input operations, geometry, EditHistory.
#pragma once

#include "RawEngine.hpp"

#include <array>
#include <cstddef>
#include <cstdint>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <string_view>
#include <variant>
#include <vector>

namespace rawengine {

// EditGraph.hpp:18-27: JSON-compatible value tree with ordered object keys.
struct EditValue {
    using Array = std::vector<EditValue>;
    using Object = std::map<std::string, EditValue>;
    std::variant<std::nullptr_t, bool, std::int64_t, double, std::string, Array, Object> data = nullptr;
    bool operator==(const EditValue&) const = default;
};

// EditGraph.hpp:29-39.
enum class EditSourceKind { DecodedBayerU16, SceneLinearRasterF32, IccRasterU16, CoverageRasterF32,
                            PremultipliedRgbaRasterF32 };
enum class EditDomain { CameraLinear, SceneLinearProPhotoD50, SceneLinearRec2020D65, SceneLinearSrgb,
                        ToneMappedUnmanaged, DisplayLinearSrgb, DisplayEncodedSrgb, DisplayEncodedIcc,
                        UnmanagedBounded, Coverage, PremultipliedSceneLinearProPhotoD50,
                        PremultipliedSceneLinearRec2020D65 };
enum class LegacyRecipeEra { ImplicitRec2020, ImplicitProPhoto };
inline constexpr std::uint32_t kLegacyRec2020ProcessingVersion = 1;
inline constexpr std::uint32_t kCurrentEditProcessingVersion = 2;

// EditGraph.hpp:41-50.
struct EditSource {
    std::string id;  // lowercase UUID
    EditSourceKind kind = EditSourceKind::SceneLinearRasterF32;
    std::optional<WorkingSpace> working_space;
    std::array<std::uint8_t, 32> content_sha256{};
    std::optional<IccProfileIdentity> icc_input;
    std::optional<RawDemosaicIdentity> demosaic;
    bool operator==(const EditSource&) const = default;
};

// EditGraph.hpp:52-67.
struct EditOperation {
    std::string id;
    std::string type_id;
    std::uint32_t schema_version = 1;
    std::uint32_t processing_version = 1;
    bool enabled = true;
    EditDomain input_domain = EditDomain::SceneLinearProPhotoD50;
    EditDomain output_domain = EditDomain::SceneLinearProPhotoD50;
    EditValue::Object parameters;
    std::map<std::string, std::string> inputs;
    std::map<std::string, std::string> masks;
    std::string blend_mode = "normal";
    double opacity = 1.0;
    EditValue::Object extra_fields;
    bool operator==(const EditOperation&) const = default;
};

// EditGraph.hpp:69-79.
struct EditManifest {
    std::uint32_t format_version = 2;
    std::uint32_t processing_version = kCurrentEditProcessingVersion;
    WorkingSpace working_space = WorkingSpace::LinearProPhotoD50;
    std::vector<EditSource> sources;
    std::vector<EditOperation> operations;
    std::string output_id;
    std::optional<IccProfileIdentity> output_profile;
    EditValue::Object extra_fields;
    bool operator==(const EditManifest&) const = default;
};

// EditGraph.hpp:81-96.  Parameters of operation types outside the render path
// are carried, not checked; those operations are rejected when a graph is built.
void validate_edit_manifest(const EditManifest& manifest);
std::string serialize_edit_manifest(const EditManifest& manifest);
EditManifest parse_edit_manifest(std::string_view json);
EditManifest snapshot_legacy_recipe(EditSource source, GraphRecipe recipe, LegacyRecipeEra era,
                                    std::string operation_id,
                                    std::optional<IccProfileIdentity> output_profile = std::nullopt);

// EditGraph.hpp:98-104 (RGB sources only).
struct BoundEditSource {
    EditSource identity;
    std::shared_ptr<const Node> node;
    Rect bounds;
};

namespace detail {
struct DeviceTile;
struct CacheAccess;
}

// EditGraph.hpp:106-145: process-local LRU cache shared across graph
// revisions.  The byte budget charges 256 bytes per entry plus the tile's
// float32 bytes; tiles that do not fit render normally without entering the
// cache.  Entries live in device memory (HBM); render() hands back host copies.
class TileCache final {
public:
    struct Stats {
        std::size_t entries = 0, used_bytes = 0, hits = 0, misses = 0;
    };
    explicit TileCache(std::size_t max_bytes);
    ~TileCache();
    TileCache(const TileCache&) = delete;
    TileCache& operator=(const TileCache&) = delete;
    // `node` renders `bounds` in the stated level's coordinates; distinct
    // mip/quality values never share an entry.
    Tile render(const Node& node, std::array<std::uint8_t, 32> signature, Rect bounds, RenderLevel level = {});
    Stats stats() const;
    void clear();

private:
    friend struct detail::CacheAccess;
    struct Key {
        std::array<std::uint8_t, 32> signature{};
        std::array<std::uint32_t, 4> bounds{};
        std::uint32_t mip = 0;
        RenderQuality quality = RenderQuality::Final;
        std::uint8_t payload = 0;
        auto operator<=>(const Key&) const = default;
    };
    struct Entry {
        std::shared_ptr<const detail::DeviceTile> tile;
        std::list<Key>::iterator recency;
        std::size_t charged_bytes = 0;
    };
    std::size_t max_bytes_ = 0, used_bytes_ = 0, hits_ = 0, misses_ = 0;
    std::uint64_t generation_ = 0;
    std::map<Key, Entry> entries_;
    std::list<Key> recency_;
    mutable std::mutex mutex_;
};

// EditGraph.hpp:147-206: immutable executable view of a manifest.  Source
// buffers and the ICC transform are owned by the caller's shared nodes and
// retained by this graph.
class ExecutableEditGraph final {
public:
    ExecutableEditGraph(EditManifest manifest, std::vector<BoundEditSource> sources,
                        std::shared_ptr<const IccDisplayTransform> display_transform = nullptr,
                        std::shared_ptr<TileCache> cache = nullptr);
    const Node& output() const { return *output_; }
    std::shared_ptr<const Node> output_handle() const { return output_; }
    bool output_is_coverage() const noexcept { return false; }
    bool output_is_rgba() const noexcept { return false; }
    Rect source_bounds() const;
    Rect output_bounds() const noexcept { return output_bounds_; }
    Rect required_source_region(Rect output) const;
    Rect required_source_region(Rect output, RenderLevel level) const;
    std::map<std::string, Rect> required_source_regions(Rect output, RenderLevel level = {}) const;
    const EditManifest& manifest() const noexcept { return manifest_; }

private:
    EditManifest manifest_;
    Rect output_bounds_;
    std::map<const Node*, std::pair<std::string, Rect>> source_nodes_;
    std::shared_ptr<const Node> output_;
};

namespace detail {
// Tile-cache signature of one operation (EditGraph.cpp:1736-1753): SHA-256 of
// the version tag, the upstream signature and the operation's canonical JSON.
// Exposed for tests.
std::array<std::uint8_t, 32> edit_operation_signature(const EditOperation& op,
                                                      const std::array<std::uint8_t, 32>& upstream);
std::array<std::uint8_t, 32> edit_source_signature(const EditSource& source);
}  // namespace detail

}  // namespace rawengine
