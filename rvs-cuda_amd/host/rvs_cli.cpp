// rvs_cli.cpp -- command-line front end with the reference's flags
// (rvs/src/view_synthesis.cpp:64-176): rvs CONFIGURATION_FILE [--cuda] [--opengl]
// [--analyzer] [--help].  The gfx950 path always runs; --cuda / --opengl are
// accepted for compatibility.  Reports total and per-frame wall time.
#include <chrono>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>

#include "rvs/pipeline.hpp"

int main(int argc, char* argv[]) {
    try {
        std::string filename, source;
        for (int i = 1; i < argc; ++i) {
            if (!std::strcmp(argv[i], "--opengl") || !std::strcmp(argv[i], "--cuda") ||
                !std::strcmp(argv[i], "--analyzer")) {
                continue;
            } else if (!std::strcmp(argv[i], "--help")) {
                filename.clear();
                break;
            } else if (!std::strcmp(argv[i], "--source") && i + 1 < argc) {
                source = argv[++i];
            } else if (filename.empty()) {
                filename = argv[i];
            } else {
                throw std::runtime_error("Too many parameters (try --help)");
            }
        }
        if (filename.empty()) throw std::runtime_error("Usage: rvs CONFIGURATION_FILE [--cuda] [--source DIR]");
        rvs::Application app(filename, source);
        std::cout << rvs_version() << std::endl;
        const auto t0 = std::chrono::steady_clock::now();
        const int frames = app.execute();
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "\nTotal time:         " << std::fixed << std::setprecision(3) << s << " sec.\n"
                  << "Av. speed:          " << frames / s << " fps.\n"
                  << "Av. time per frame: " << s / frames << " sec." << std::endl;
        return 0;
    } catch (std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
