// rawengine render-path API, MI355X build.
//
// Source-compatible subset of the reference's public header for the tiled
// render path (RawEngine.hpp:26-710 of tylermcm/librawops): the data model,
// the Node interface, the RAW/raster source nodes, the point nodes of the
// render chain, ImageGraph, RenderRequest and Renderer::render_tiles /
// render_image / render_roi.  Behind it, chains of these nodes are compiled
// into fused gfx950 kernels (include/rawops_hip.h); a node type this library
// does not know still renders through its own Node::render_level, exactly as
// the reference's tile loop would call it.
//
// Additive extensions (no counterpart in the reference):
//   * ImageGraph(RawImage, GraphRecipe, RawDemosaicIdentity, bool linear_output)
//     selects Menon and/or ends the chain after camera calibration;
//   * Renderer::render_device(...) renders into caller-owned device memory.
//
// The ICC hook (IccDisplayTransform / IccDisplayNode, RawEngine.hpp:595-619)
// keeps the reference's host callback contract: the transform runs on the
// host over the device-rendered tile, because a colour-management engine
// (LittleCMS in the reference) is host code.
#pragma once

#include <array>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rawengine {

enum class BayerPattern { RGGB, BGGR, GRBG, GBRG };
enum class WorkingSpace { LinearProPhotoD50, LinearRec2020D65 };

struct Rect {
    std::uint32_t x = 0, y = 0, width = 0, height = 0;
};

enum class RenderQuality : std::uint8_t { Preview = 0, Final = 1 };

struct RenderLevel {
    std::uint32_t mip = 0;
    RenderQuality quality = RenderQuality::Final;
};

struct RenderRequest {
    Rect viewport;
    std::uint32_t tile_size = 256;
    RenderLevel level;
    RenderRequest() = default;
    explicit constexpr RenderRequest(Rect requested_viewport, std::uint32_t requested_tile_size = 256,
                                     RenderLevel requested_level = {}) noexcept
        : viewport(requested_viewport), tile_size(requested_tile_size), level(requested_level) {}
};

struct RawMetadata {
    std::uint32_t width = 0, height = 0;
    std::uint32_t row_stride_samples = 0;  // 0 means tightly packed width.
    BayerPattern pattern = BayerPattern::RGGB;
    std::uint8_t cfa_phase_x = 0, cfa_phase_y = 0;
    Rect active_area;  // all zero means the full sensor
    std::array<std::uint16_t, 4> black_levels{0, 0, 0, 0};
    std::array<std::uint16_t, 4> white_levels{65535, 65535, 65535, 65535};
};

namespace detail {
struct DeviceSource;
}

class RawImage;
class RasterImage;
// Version-1 canonical source digests (SourceFingerprint.cpp:61-104): SHA-256
// over the rendering-relevant metadata and the visible samples in row-major
// order, little-endian; row padding does not enter.
std::array<std::uint8_t, 32> fingerprint_raw_source(const RawImage& image);
std::array<std::uint8_t, 32> fingerprint_raster_source(const RasterImage& image);

// Decoded uint16 Bayer samples.  Copies share immutable storage; the samples
// are uploaded to HBM once, on first render, and stay resident.
class RawImage {
public:
    RawImage(RawMetadata metadata, std::vector<std::uint16_t> samples);
    RawImage(std::uint32_t width, std::uint32_t height, std::vector<std::uint16_t> samples,
             BayerPattern pattern = BayerPattern::RGGB, std::uint16_t black_level = 0,
             std::uint16_t white_level = 65535);
    const RawMetadata& metadata() const noexcept { return metadata_; }
    const std::vector<std::uint16_t>& samples() const noexcept { return *bayer_; }
    std::uint32_t width() const noexcept { return metadata_.width; }
    std::uint32_t height() const noexcept { return metadata_.height; }
    std::array<std::uint8_t, 32> fingerprint() const;  // computed once, shared by copies
    const detail::DeviceSource& device() const;  // uploads on first use
private:
    RawMetadata metadata_;
    std::shared_ptr<const std::vector<std::uint16_t>> bayer_;
    std::shared_ptr<std::once_flag> fingerprint_once_ = std::make_shared<std::once_flag>();
    std::shared_ptr<std::array<std::uint8_t, 32>> fingerprint_ = std::make_shared<std::array<std::uint8_t, 32>>();
    std::shared_ptr<std::once_flag> upload_once_ = std::make_shared<std::once_flag>();
    std::shared_ptr<std::shared_ptr<detail::DeviceSource>> device_ =
        std::make_shared<std::shared_ptr<detail::DeviceSource>>();
};

struct RasterMetadata {
    std::uint32_t width = 0, height = 0;
    std::uint32_t row_stride_pixels = 0;
    WorkingSpace working_space = WorkingSpace::LinearProPhotoD50;
};

class RasterImage {
public:
    RasterImage(RasterMetadata metadata, std::vector<float> pixels);
    const RasterMetadata& metadata() const noexcept { return metadata_; }
    const std::vector<float>& pixels() const noexcept { return *pixels_; }
    std::uint32_t width() const noexcept { return metadata_.width; }
    std::uint32_t height() const noexcept { return metadata_.height; }
    std::array<std::uint8_t, 32> fingerprint() const;
    const detail::DeviceSource& device() const;
private:
    RasterMetadata metadata_;
    std::shared_ptr<const std::vector<float>> pixels_;
    std::shared_ptr<std::once_flag> fingerprint_once_ = std::make_shared<std::once_flag>();
    std::shared_ptr<std::array<std::uint8_t, 32>> fingerprint_ = std::make_shared<std::array<std::uint8_t, 32>>();
    std::shared_ptr<std::once_flag> upload_once_ = std::make_shared<std::once_flag>();
    std::shared_ptr<std::shared_ptr<detail::DeviceSource>> device_ =
        std::make_shared<std::shared_ptr<detail::DeviceSource>>();
};

struct CameraColorTransform {
    std::array<double, 9> camera_to_xyz_d50{};
    WorkingSpace target = WorkingSpace::LinearProPhotoD50;
};

enum class OutputMode { LegacyBounded, SrgbPreview, IccDisplay };

struct GraphRecipe {
    float red_gain = 1.0f, green_gain = 1.0f, blue_gain = 1.0f;
    float exposure_stops = 0.0f;
    float tone_shoulder = 0.25f;
    float tone_gamma = 1.0f;
    std::optional<CameraColorTransform> camera_color;
    OutputMode output_mode = OutputMode::LegacyBounded;
};

enum class PixelFormat { RGBFloat32 };
enum class PixelDomain { CameraLinearRGB, SceneLinearRGB, ToneMappedUnmanagedRGB,
                         BoundedUnmanagedRGB, DisplayLinearRGB, DisplayEncodedRGB };
enum class ColorPrimaries { CameraNative, ProPhoto, Rec2020, SRGB, ICCProfile };
enum class WhitePoint { Unspecified, D50, D65 };
enum class TransferFunction { Linear, CustomTone, SRGB, ICCProfile };
enum class ReferenceState { CameraReferred, SceneReferred, DisplayReferred, Unspecified };
enum class AlphaMode { None };

struct ImageDescriptor {
    PixelFormat format = PixelFormat::RGBFloat32;
    PixelDomain domain = PixelDomain::CameraLinearRGB;
    ColorPrimaries primaries = ColorPrimaries::CameraNative;
    WhitePoint white_point = WhitePoint::Unspecified;
    TransferFunction transfer = TransferFunction::Linear;
    ReferenceState reference = ReferenceState::CameraReferred;
    AlphaMode alpha = AlphaMode::None;
    std::array<std::uint8_t, 32> profile_sha256{};

    static constexpr ImageDescriptor camera_linear() noexcept { return {}; }
    static constexpr ImageDescriptor scene_linear(WorkingSpace space) noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::SceneLinearRGB,
                space == WorkingSpace::LinearProPhotoD50 ? ColorPrimaries::ProPhoto : ColorPrimaries::Rec2020,
                space == WorkingSpace::LinearProPhotoD50 ? WhitePoint::D50 : WhitePoint::D65,
                TransferFunction::Linear, ReferenceState::SceneReferred, AlphaMode::None};
    }
    static constexpr ImageDescriptor linear_srgb() noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::SceneLinearRGB, ColorPrimaries::SRGB, WhitePoint::D65,
                TransferFunction::Linear, ReferenceState::SceneReferred, AlphaMode::None};
    }
    static constexpr ImageDescriptor tone_mapped() noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::ToneMappedUnmanagedRGB, ColorPrimaries::CameraNative,
                WhitePoint::Unspecified, TransferFunction::CustomTone, ReferenceState::Unspecified, AlphaMode::None};
    }
    static constexpr ImageDescriptor bounded_output() noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::BoundedUnmanagedRGB, ColorPrimaries::CameraNative,
                WhitePoint::Unspecified, TransferFunction::CustomTone, ReferenceState::Unspecified, AlphaMode::None};
    }
    static constexpr ImageDescriptor srgb_output() noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::DisplayEncodedRGB, ColorPrimaries::SRGB, WhitePoint::D65,
                TransferFunction::SRGB, ReferenceState::DisplayReferred, AlphaMode::None};
    }
    static constexpr ImageDescriptor display_linear_srgb() noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::DisplayLinearRGB, ColorPrimaries::SRGB, WhitePoint::D65,
                TransferFunction::Linear, ReferenceState::DisplayReferred, AlphaMode::None};
    }
    static constexpr ImageDescriptor icc_display(std::array<std::uint8_t, 32> digest) noexcept {
        return {PixelFormat::RGBFloat32, PixelDomain::DisplayEncodedRGB, ColorPrimaries::ICCProfile,
                WhitePoint::Unspecified, TransferFunction::ICCProfile, ReferenceState::DisplayReferred,
                AlphaMode::None, digest};
    }
    bool operator==(const ImageDescriptor&) const = default;
};

struct Tile {
    Rect bounds;
    std::vector<float> rgb;  // interleaved float32 RGB, row-major
    ImageDescriptor descriptor = ImageDescriptor::camera_linear();
};

// RawEngine.hpp:212-219: exact output/input profile and conversion policy.
struct IccProfileIdentity {
    std::array<std::uint8_t, 32> profile_sha256{};
    std::string intent = "relative_colorimetric";
    bool black_point_compensation = false;
    std::string engine = "lcms2-core";
    std::string engine_version = "2.19.1";
    bool operator==(const IccProfileIdentity&) const = default;
};

struct RawDemosaicIdentity {
    std::string algorithm = "rawengine.bilinear";
    std::uint32_t processing_version = 1;
    bool operator==(const RawDemosaicIdentity&) const = default;
};
void validate_raw_demosaic(const RawDemosaicIdentity& identity);

class Node {
public:
    virtual ~Node() = default;
    virtual Tile render(Rect bounds) const = 0;
    virtual Tile render_level(Rect bounds, RenderLevel level) const {
        if (level.mip == 0 && level.quality == RenderQuality::Final) return render(bounds);
        throw std::invalid_argument("node does not support this render level");
    }
    virtual bool supports_level(RenderLevel level) const noexcept {
        return level.mip == 0 && level.quality == RenderQuality::Final;
    }
    virtual ImageDescriptor output_descriptor() const noexcept = 0;
    virtual const Node* input_node() const noexcept { return nullptr; }
    virtual RenderLevel input_level(RenderLevel output_level) const { return output_level; }
    virtual Rect input_region(Rect output, Rect source_bounds) const {
        (void)source_bounds;
        return output;
    }
    virtual Rect input_region_level(Rect output, Rect source_bounds, RenderLevel level) const {
        if (level.mip == 0 && level.quality == RenderQuality::Final) return input_region(output, source_bounds);
        throw std::invalid_argument("node has no mapping for this render level");
    }
    virtual std::optional<IccProfileIdentity> input_icc_identity() const { return std::nullopt; }
    virtual std::optional<std::array<std::uint8_t, 32>> source_fingerprint() const { return std::nullopt; }
    virtual std::optional<Rect> source_bounds() const { return std::nullopt; }
    virtual std::optional<RawDemosaicIdentity> raw_demosaic_identity() const { return std::nullopt; }
};

// Built-in nodes render through the device: render(bounds) compiles the chain
// ending at this node and runs it on the GPU.
class DeviceNode : public Node {
public:
    Tile render(Rect bounds) const override;
    Tile render_level(Rect bounds, RenderLevel level) const override;
};

class RawUnpackNode final : public DeviceNode {
public:
    explicit RawUnpackNode(RawImage image);
    RawUnpackNode(RawImage image, RawDemosaicIdentity demosaic);
    Rect input_region(Rect output, Rect source_bounds) const override;
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::camera_linear(); }
    std::optional<Rect> source_bounds() const override { return image_.metadata().active_area; }
    std::optional<std::array<std::uint8_t, 32>> source_fingerprint() const override {
        return fingerprint_raw_source(image_);
    }
    std::optional<RawDemosaicIdentity> raw_demosaic_identity() const override { return demosaic_; }
    const RawImage& image() const noexcept { return image_; }
private:
    RawImage image_;
    RawDemosaicIdentity demosaic_;
};

class RasterSourceNode final : public DeviceNode {
public:
    explicit RasterSourceNode(RasterImage image);
    bool supports_level(RenderLevel level) const noexcept override {
        return (level.mip == 0 && (level.quality == RenderQuality::Final || level.quality == RenderQuality::Preview)) ||
               (level.mip >= 1 && level.mip <= 2 && level.quality == RenderQuality::Preview);
    }
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level)) throw std::invalid_argument("raster source has no mapping for this render level");
        return output;
    }
    ImageDescriptor output_descriptor() const noexcept override {
        return ImageDescriptor::scene_linear(image_.metadata().working_space);
    }
    std::optional<Rect> source_bounds() const override { return Rect{0, 0, image_.width(), image_.height()}; }
    std::optional<std::array<std::uint8_t, 32>> source_fingerprint() const override {
        return fingerprint_raster_source(image_);
    }
    const RasterImage& image() const noexcept { return image_; }
private:
    RasterImage image_;
};

class WhiteBalanceNode final : public DeviceNode {
public:
    WhiteBalanceNode(std::shared_ptr<const Node> input, float red, float green, float blue);
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::camera_linear(); }
    const float* gains() const noexcept { return gains_; }
private:
    std::shared_ptr<const Node> input_;
    float gains_[3];
};

class ExposureNode final : public DeviceNode {
public:
    ExposureNode(std::shared_ptr<const Node> input, float stops);
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level)) throw std::invalid_argument("exposure has no mapping for this render level");
        return output;
    }
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
    float multiplier() const noexcept { return multiplier_; }
private:
    std::shared_ptr<const Node> input_;
    float multiplier_;
    ImageDescriptor descriptor_;
};

class CameraToWorkingNode final : public DeviceNode {
public:
    CameraToWorkingNode(std::shared_ptr<const Node> input, CameraColorTransform transform,
                        Rect native_input_bounds = {});
    bool supports_level(RenderLevel level) const noexcept override;
    RenderLevel input_level(RenderLevel level) const override;
    Rect input_region_level(Rect output, Rect input_bounds, RenderLevel level) const override;
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
    const std::array<float, 9>& matrix() const noexcept { return matrix_; }
    Rect native_input_bounds() const noexcept { return native_input_bounds_; }
private:
    std::shared_ptr<const Node> input_;
    std::array<float, 9> matrix_{};
    ImageDescriptor descriptor_;
    Rect native_input_bounds_;
};

class WorkingSpaceConvertNode final : public DeviceNode {
public:
    WorkingSpaceConvertNode(std::shared_ptr<const Node> input, WorkingSpace target);
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level))
            throw std::invalid_argument("working-space conversion has no mapping for this render level");
        return output;
    }
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
    const std::array<float, 9>& matrix() const noexcept { return matrix_; }
    bool identity() const noexcept { return identity_; }
private:
    std::shared_ptr<const Node> input_;
    std::array<float, 9> matrix_{};
    ImageDescriptor descriptor_;
    bool identity_ = false;
};

class WorkingToSrgbNode final : public DeviceNode {
public:
    explicit WorkingToSrgbNode(std::shared_ptr<const Node> input);
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level)) throw std::invalid_argument("sRGB preview stage has no mapping for this render level");
        return output;
    }
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::linear_srgb(); }
    const std::array<float, 9>& matrix() const noexcept { return matrix_; }
private:
    std::shared_ptr<const Node> input_;
    std::array<float, 9> matrix_{};
};

class ToneCurveNode final : public DeviceNode {
public:
    ToneCurveNode(std::shared_ptr<const Node> input, float shoulder, float gamma);
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level)) throw std::invalid_argument("sRGB preview stage has no mapping for this render level");
        return output;
    }
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
    float shoulder() const noexcept { return shoulder_; }
    float inverse_gamma() const noexcept { return inverse_gamma_; }
private:
    std::shared_ptr<const Node> input_;
    float shoulder_, inverse_gamma_;
    ImageDescriptor descriptor_;
};

class OutputClipNode final : public DeviceNode {
public:
    explicit OutputClipNode(std::shared_ptr<const Node> input);
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
private:
    std::shared_ptr<const Node> input_;
    ImageDescriptor descriptor_;
};

class SrgbEncodeNode final : public DeviceNode {
public:
    explicit SrgbEncodeNode(std::shared_ptr<const Node> input);
    bool supports_level(RenderLevel level) const noexcept override;
    Rect input_region_level(Rect output, Rect, RenderLevel level) const override {
        if (!supports_level(level)) throw std::invalid_argument("sRGB preview stage has no mapping for this render level");
        return output;
    }
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return ImageDescriptor::srgb_output(); }
private:
    std::shared_ptr<const Node> input_;
};

// RawEngine.hpp:595-606: output ICC transform from display-linear sRGB to
// bounded, profile-encoded float32 RGB (thread-safe for concurrent calls).
class IccDisplayTransform {
public:
    virtual ~IccDisplayTransform() = default;
    virtual std::array<std::uint8_t, 32> profile_sha256() const noexcept = 0;
    virtual std::optional<IccProfileIdentity> output_icc_identity() const { return std::nullopt; }
    virtual void apply(float* interleaved_rgb, std::size_t pixels) const = 0;
};

// RawEngine.hpp:608-619, RawEngine.cpp:1279-1304.  Renders its input on the
// device, then runs the transform over the host tile; non-finite input raises
// domain_error, as does output outside [0, 1].
class IccDisplayNode final : public Node {
public:
    IccDisplayNode(std::shared_ptr<const Node> input, std::shared_ptr<const IccDisplayTransform> transform);
    Tile render(Rect bounds) const override;
    const Node* input_node() const noexcept override { return input_.get(); }
    ImageDescriptor output_descriptor() const noexcept override { return descriptor_; }
private:
    std::shared_ptr<const Node> input_;
    std::shared_ptr<const IccDisplayTransform> transform_;
    ImageDescriptor descriptor_;
};

class ImageGraph final {
public:
    ImageGraph(RawImage image, GraphRecipe recipe = {},
               std::shared_ptr<const IccDisplayTransform> display_transform = nullptr);
    // Extension: demosaic selection and linear (post-calibration) output.
    ImageGraph(RawImage image, GraphRecipe recipe, RawDemosaicIdentity demosaic, bool linear_output = false,
               std::shared_ptr<const IccDisplayTransform> display_transform = nullptr);
    ImageGraph(RasterImage image, GraphRecipe recipe = {},
               std::shared_ptr<const IccDisplayTransform> display_transform = nullptr);
    ImageGraph(std::shared_ptr<const Node> scene_linear_source, Rect source_bounds, GraphRecipe recipe = {},
               std::shared_ptr<const IccDisplayTransform> display_transform = nullptr);
    const RawImage& image() const;
    Rect source_bounds() const noexcept { return source_bounds_; }
    const GraphRecipe& recipe() const noexcept { return recipe_; }
    const Node& output() const noexcept { return *output_; }
private:
    void finish(std::shared_ptr<const Node> linear, bool linear_output,
                std::shared_ptr<const IccDisplayTransform> display_transform);
    std::optional<RawImage> raw_image_;
    Rect source_bounds_;
    GraphRecipe recipe_;
    std::shared_ptr<const Node> output_;
};

class ExecutableEditGraph;

class RenderCancelled final : public std::runtime_error {
public:
    RenderCancelled() : std::runtime_error("render cancelled") {}
};

class CancellationToken final {
public:
    void cancel() noexcept { cancelled_.store(true, std::memory_order_relaxed); }
    bool is_cancelled() const noexcept { return cancelled_.load(std::memory_order_relaxed); }
private:
    std::atomic<bool> cancelled_{false};
};

class Renderer final {
public:
    using TileCallback = std::function<void(const Tile&)>;
    void render_tiles(const ImageGraph& graph, RenderRequest request, const TileCallback& callback,
                      const CancellationToken* cancellation = nullptr) const;
    void render_tiles(const ExecutableEditGraph& graph, RenderRequest request, const TileCallback& callback,
                      const CancellationToken* cancellation = nullptr) const;
    void render_tiles(const Node& output, Rect source_bounds, RenderRequest request, const TileCallback& callback,
                      const CancellationToken* cancellation = nullptr) const;
    void render_tiles(const ImageGraph& graph, Rect viewport, const TileCallback& callback,
                      std::uint32_t tile_size = 256, const CancellationToken* cancellation = nullptr) const;
    void render_tiles(const ExecutableEditGraph& graph, Rect viewport, const TileCallback& callback,
                      std::uint32_t tile_size = 256, const CancellationToken* cancellation = nullptr) const;
    void render_tiles(const Node& output, Rect source_bounds, Rect viewport, const TileCallback& callback,
                      std::uint32_t tile_size = 256, const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const ImageGraph& graph, RenderRequest request,
                      const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const ExecutableEditGraph& graph, RenderRequest request,
                      const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const Node& output, Rect source_bounds, RenderRequest request,
                      const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const ImageGraph& graph, Rect viewport, std::uint32_t tile_size = 256,
                      const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const ExecutableEditGraph& graph, Rect viewport, std::uint32_t tile_size = 256,
                      const CancellationToken* cancellation = nullptr) const;
    Tile render_image(const Node& output, Rect source_bounds, Rect viewport, std::uint32_t tile_size = 256,
                      const CancellationToken* cancellation = nullptr) const;
    std::vector<float> render_roi(const ImageGraph& graph, Rect viewport, std::uint32_t tile_size = 256,
                                  const CancellationToken* cancellation = nullptr) const;
    std::vector<float> render_roi(const ExecutableEditGraph& graph, Rect viewport, std::uint32_t tile_size = 256,
                                  const CancellationToken* cancellation = nullptr) const;
    // Extension: render into caller-owned device memory (float32 RGB, row
    // stride = viewport width) on `stream` (a hipStream_t), asynchronously.
    // Non-finite output errors are reported by the next synchronous call.
    void render_device(const Node& output, Rect source_bounds, RenderRequest request, float* device_rgb,
                       void* stream = nullptr) const;
};

}  // namespace rawengine
