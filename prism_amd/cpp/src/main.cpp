// main.cpp -- the `prism` command line (flags of the reference's
// src/utils/parameters.cu; see usage()).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>

#include "prism/prism.hpp"

int main(int argc, char** argv) {
    prism_context config;
    try {
        parse_argv(&config, argc, argv);
    } catch (const std::invalid_argument& e) {
        std::fprintf(stderr, "prism: %s\n", e.what());
        return EXIT_FAILURE;
    }
    initialize(&config);
    release_context(&config);
    return 0;
}
