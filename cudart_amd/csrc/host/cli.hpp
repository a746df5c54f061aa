// Command-line front end shared by the product renderer (`prt`, HIP) and the CPU
// oracle's renderer (`prt_cpu`): option parsing, scene loading and the render /
// save loop of src/main.cpp and TerminalLauncher (src/context/terminal_launcher.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "prt.h"

namespace prt::cli {

struct Options {
    int samples = 10;                 // -s (main.cpp:26)
    int width = 640;                  // -w (main.cpp:27)
    int height = 480;                 // -h (main.cpp:28)
    std::string scene = "cornell";    // -f (main.cpp:30)
    std::string out;                  // -o; default <scene>.png (main.cpp:29,125)
    std::string env = "../.env";      // --env (main.cpp:36)
    std::string pfm;                  // --pfm: float radiance as well
    std::vector<std::pair<std::string, float>> emit;  // --emit name=Ka
    std::string integrator = "megakernel";
    uint32_t max_depth = 6;
    uint32_t seed = 0x5EED;
    int device = 0;
    bool cpu = false;                 // --cpu (main.cpp:101-104)
    bool progressive = false;         // --x: rewrite the image after every sample pass
    // resolved paths
    std::string obj_path, camera_path, textures_path;
};

// The per-backend part: build the scene, render one sample of every pixel,
// read the 8-bit image and the mean radiance.
class Backend {
public:
    virtual ~Backend() = default;
    virtual const char* name() const = 0;
    virtual int create(const prt_scene_desc& desc, const prt_camera& cam, const Options& o) = 0;
    virtual int render_one() = 0;  // renderOneSamplePerPixel
    virtual int image_bytes(uint8_t* rgba) = 0;
    virtual int radiance(float* rgb) = 0;
    virtual uint32_t samples_rendered() const = 0;
};

// Parses argv. Returns 0 to run, 1 after printing usage, 2 on an invalid value.
int parse(int argc, char** argv, Options& o);

// Runs main.cpp's sequence with the backend: settings, scene load (OBJ + MTL,
// camera JSON or a default camera framing the scene), render loop, PNG (and PFM).
int run(int argc, char** argv, Backend& be);

}  // namespace prt::cli
