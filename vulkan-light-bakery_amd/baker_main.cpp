// baker -- command-line driver with the reference's shape (src/baker/main.cpp:
// 8-44): construct LightBaker, bake(), serialize(); exceptions map to
// EXIT_FAILURE.  Extra options select the library's other bake forms.
//
//   baker <scene.gltf|scene.glb|room|terrain|city|cube|spheres[:seed]>
//         [--grid N] [--rays R | --equirect W H] [--lambert] [--order2]
//         [--bf16|--fp16]
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <string>

#include "vlb/light_baker.hpp"

int main(int argc, char** argv)
{
    try {
        if (argc < 2) throw std::runtime_error("Select scene to bake.");
        std::string sceneName = argv[1];
        vlb::LightBaker::Settings s;
        for (int i = 2; i < argc; ++i) {
            std::string a = argv[i];
            if (a == "--grid" && i + 1 < argc) {
                uint32_t n = (uint32_t)std::stoul(argv[++i]);
                s.probesCount3D = {n, n, n};
            } else if (a == "--rays" && i + 1 < argc) {
                s.dirMode = VLB_DIRS_FIBONACCI;
                s.rays = (uint32_t)std::stoul(argv[++i]);
            } else if (a == "--equirect" && i + 2 < argc) {
                s.dirMode = VLB_DIRS_EQUIRECT;
                s.equirectW = (uint32_t)std::stoul(argv[++i]);
                s.equirectH = (uint32_t)std::stoul(argv[++i]);
            } else if (a == "--lambert") {
                s.shadeMode = VLB_SHADE_LAMBERT;
                s.shConstants = VLB_SH_EXACT;
                s.lights.push_back(vlb_light{VLB_LIGHT_DIRECTIONAL, {-0.3f, -0.9f, -0.3f}, {3.0f, 3.0f, 3.0f}});
            } else if (a == "--order2") {
                s.shOrder = 2;
            } else if (a == "--bf16") {
                s.precision = VLB_PREC_BF16;
            } else if (a == "--fp16") {
                s.precision = VLB_PREC_FP16;
            } else {
                throw std::runtime_error("unknown option " + a);
            }
        }
        vlb::LightBaker baker{sceneName, 0, s};
        const auto t0 = std::chrono::steady_clock::now();
        baker.bake();
        const auto t1 = std::chrono::steady_clock::now();
        baker.serialize();
        const double ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        std::cout << "baked " << baker.probePositions().size() << " probes in " << ms << " ms -> "
                  << baker.outputName() << "\n";
    } catch (std::exception& error) {
        std::cerr << "std::exception: " << error.what() << "\n";
        return EXIT_FAILURE;
    } catch (...) {
        std::cerr << "unknown error\n";
        return EXIT_FAILURE;
    }
    std::cout << "exiting...\n";
    return EXIT_SUCCESS;
}
