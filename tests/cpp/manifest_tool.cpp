// Command-line probe of this library's host-side edit-manifest code (no GPU
// calls), driven by tests/test_edit_manifest.py:
//   manifest_tool canon PATH                       canonical JSON of a manifest
//   manifest_tool fingerprint-raw PATH W H STRIDE PATTERN PX PY AX AY AW AH B0 B1 B2 B3 W0 W1 W2 W3
//   manifest_tool fingerprint-raster PATH W H STRIDE SPACE
//   manifest_tool legacy RAW DIGEST SRC OP ERA G0 G1 G2 STOPS SHOULDER GAMMA MODE HASCAM C0..C8 TARGET
//   manifest_tool signatures PATH                  one line per source / enabled operation:
//                                                  id, upstream signature, own signature (hex)
// Errors print "ERROR <code> <message>" (1 invalid_argument, 2 out_of_range,
// 4 length_error, 7 other) and exit 0.
#include <EditGraph.hpp>

#include <cstdio>
#include <cstring>
#include <map>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <sstream>
#include <string>

using namespace rawengine;

static std::string slurp(const char* path) {
    std::ifstream f(path, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

static std::string hex(const std::array<std::uint8_t, 32>& d) {
    static const char* h = "0123456789abcdef";
    std::string s;
    for (auto b : d) { s += h[b >> 4]; s += h[b & 15]; }
    return s;
}

static std::array<std::uint8_t, 32> unhex(const std::string& s) {
    std::array<std::uint8_t, 32> d{};
    for (int i = 0; i < 32; ++i) d[i] = std::uint8_t(std::stoul(s.substr(2 * i, 2), nullptr, 16));
    return d;
}

static int body(int argc, char** argv) {
    const std::string cmd = argv[1];
    auto u = [&](int i) { return std::uint32_t(std::stoul(argv[i])); };
    if (cmd == "canon") {
        std::fputs(serialize_edit_manifest(parse_edit_manifest(slurp(argv[2]))).c_str(), stdout);
    } else if (cmd == "fingerprint-raw") {
        const std::string raw = slurp(argv[2]);
        std::vector<std::uint16_t> s(raw.size() / 2);
        std::memcpy(s.data(), raw.data(), s.size() * 2);
        RawMetadata m;
        m.width = u(3); m.height = u(4); m.row_stride_samples = u(5);
        m.pattern = BayerPattern(u(6)); m.cfa_phase_x = std::uint8_t(u(7)); m.cfa_phase_y = std::uint8_t(u(8));
        m.active_area = {u(9), u(10), u(11), u(12)};
        for (int i = 0; i < 4; ++i) { m.black_levels[i] = std::uint16_t(u(13 + i)); m.white_levels[i] = std::uint16_t(u(17 + i)); }
        const RawImage image(m, std::move(s));
        const RawImage copy = image;  // copies share the digest
        std::printf("%s %s", hex(fingerprint_raw_source(image)).c_str(), hex(copy.fingerprint()).c_str());
    } else if (cmd == "fingerprint-raster") {
        const std::string raw = slurp(argv[2]);
        std::vector<float> px(raw.size() / 4);
        std::memcpy(px.data(), raw.data(), px.size() * 4);
        const RasterImage image({u(3), u(4), u(5), WorkingSpace(u(6))}, std::move(px));
        std::printf("%s", hex(fingerprint_raster_source(image)).c_str());
    } else if (cmd == "legacy") {
        EditSource src;
        const bool raw = u(2) != 0;
        src.kind = raw ? EditSourceKind::DecodedBayerU16 : EditSourceKind::SceneLinearRasterF32;
        src.content_sha256 = unhex(argv[3]);
        src.id = argv[4];
        GraphRecipe r;
        r.red_gain = std::stof(argv[7]); r.green_gain = std::stof(argv[8]); r.blue_gain = std::stof(argv[9]);
        r.exposure_stops = std::stof(argv[10]); r.tone_shoulder = std::stof(argv[11]); r.tone_gamma = std::stof(argv[12]);
        r.output_mode = OutputMode(u(13));
        const auto target = WorkingSpace(u(24));
        if (!raw) src.working_space = target;
        if (u(14)) {
            CameraColorTransform t;
            for (int i = 0; i < 9; ++i) t.camera_to_xyz_d50[i] = std::stod(argv[15 + i]);
            t.target = target;
            r.camera_color = t;
        }
        std::fputs(serialize_edit_manifest(snapshot_legacy_recipe(src, r, LegacyRecipeEra(u(6)), argv[5])).c_str(),
                   stdout);
    } else if (cmd == "signatures") {
        const EditManifest m = parse_edit_manifest(slurp(argv[2]));
        std::map<std::string, std::array<std::uint8_t, 32>> sig;
        for (const auto& s : m.sources) {
            sig[s.id] = detail::edit_source_signature(s);
            std::printf("%s - %s\n", s.id.c_str(), hex(sig[s.id]).c_str());
        }
        // Chains in manifest order are resolved by repeated passes.
        for (size_t pass = 0; pass < m.operations.size(); ++pass)
            for (const auto& op : m.operations) {
                if (sig.count(op.id) || !op.inputs.count("image") || !sig.count(op.inputs.at("image"))) continue;
                const auto up = sig[op.inputs.at("image")];
                sig[op.id] = op.enabled ? detail::edit_operation_signature(op, up) : up;
                std::printf("%s %s %s\n", op.id.c_str(), hex(up).c_str(), hex(sig[op.id]).c_str());
            }
    } else {
        std::fprintf(stderr, "unknown command\n");
        return 2;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    try {
        return body(argc, argv);
    } catch (const std::invalid_argument& e) { std::printf("ERROR 1 %s", e.what()); }
    catch (const std::out_of_range& e) { std::printf("ERROR 2 %s", e.what()); }
    catch (const std::length_error& e) { std::printf("ERROR 4 %s", e.what()); }
    catch (const std::exception& e) { std::printf("ERROR 7 %s", e.what()); }
    return 0;
}
