// parameters.hpp -- run configuration and command line of the PRISM API.
//
// Same structure, members and defaults as the reference's
// include/utils/parameters.hpp:8-56, so drivers written against it carry
// over.  begin/end are device int32[4] (planes [begin, end) of each level
// already retrieved), allocated by parse_argv like the reference
// (parameters.cu:229-233) and released by release_context().
#pragma once

#include <string>
#include <vector>

#include "err.hpp"
#include "traits.hpp"

struct interpolation_parameters {
    double alpha{1.75};
    double beta{4.0};
    bool use_md[4];
    bool use_natural[4];
    bool reverse[4];
    uint8_t auto_tuning{3};
    interpolation_parameters()
        : use_md{true, true, true, true}, use_natural{false, false, false, false},
          reverse{false, false, false, false} {}
};

typedef struct prism_context {
    double eb{0.0};
    double rel_eb{0.0};
    double pw_rel_eb{0.0};
    uint32_t radius{0};
    int ndim{-1};
    uint32_t x{1}, y{1}, z{1}, w{1};
    uint32_t size{1};
    bool isComp{false};
    bool isDecomp{false};
    prism_dtype dtype{F4};
    prism_mode error_mode{REL};
    std::string oriFilePath;
    std::string cmpFilePath;
    std::string decFilePath;
    int blocks_num{1};
    int levels{15};
    int bitgroups{8};
    int compMode{0};
    bool report_time{false};
    bool report_cr{false};
    size_t loaded_size{0};
    btype bt{SM};
    std::vector<double> target_ebs;
    interpolation_parameters intp_param;
    int* begin{nullptr};
    int* end{nullptr};
} context;

void usage();
// Parses the reference's flags (parameters.cu:65-234).  Throws
// std::invalid_argument for missing dimensions, paths or error bounds, like
// the reference; exits with EXIT_FAILURE on an unknown option.
void parse_argv(prism_context* config, int argc, char** argv);
// Parses without allocating begin/end (used by tools that only inspect the
// command line).
void parse_argv_host(prism_context* config, int argc, char** argv);
void release_context(prism_context* config);
