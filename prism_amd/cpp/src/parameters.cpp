// parameters.cpp -- the PRISM command line (flags, defaults and file-name
// rules of the reference's src/utils/parameters.cu:9-234).
#include "prism/parameters.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>

void usage() {
    std::cout << R"(
Options:
  -i <path> : Path to binary input file containing the original data
  -x <path> : Path to decompressed binary output file (optional)
  -z <path> : Path to compressed binary output/input file (optional)
  -f : single precision (float type)
  -d : double precision (double type)
  -1 <nx> : Dimensions for 1D array a[nx]
  -2 <nx> <ny> : Dimensions for 2D array a[ny][nx]
  -3 <nx> <ny> <nz> : Dimensions for 3D array a[nz][ny][nx]
  -A [errorBound] : Absolute error bound
  -R [errorBound] : Relative error bound
  --report [time,cr] : output the compression ratio/time
  -prog : Enable progressive compression/decompression
  -errors -<nums> <error_1> <error_2> ... : Perform progressive decompression multiple times with gradually decreasing error bounds
Examples:
  prism -i [oriFilePath] -f -3 [nx] [ny] [nz] -R [errorBound] -z [cmpFilePath] -x [decFilePath] --report time,cr
  prism -i [oriFilePath] -d -3 [nx] [ny] [nz] -R [errorBound] -z [cmpFilePath] -x [decFilePath] -prog -errors -2 1e-1 1e-2
)" << std::endl;
}

namespace {

// file-name defaults and required arguments (parameters.cu:33-63)
void check(prism_context* c) {
    if (c->ndim == -1) throw std::invalid_argument("dim must be specified.");
    if (c->report_cr && c->oriFilePath.empty()) throw std::invalid_argument("input file path must be specified.");
    if (c->isComp) {
        if (c->oriFilePath.empty()) throw std::invalid_argument("input file path must be specified.");
        if (c->cmpFilePath.empty()) c->cmpFilePath = c->oriFilePath + ".prisma";
    }
    if (c->isDecomp) {
        if (c->oriFilePath.empty() && c->cmpFilePath.empty())
            throw std::invalid_argument("compressed file path must be specified.");
        if (c->decFilePath.empty()) {
            if (!c->cmpFilePath.empty()) {
                c->decFilePath = c->cmpFilePath + ".out";
            } else {
                c->cmpFilePath = c->oriFilePath + ".prisma";
                c->decFilePath = c->oriFilePath + ".prisma.out";
            }
        }
    }
    if (!c->isComp && !c->isDecomp) {
        std::printf("Please set the compression/decompression\n");
        usage();
    }
}

bool is(const char* a, const char* b) { return std::strcmp(a, b) == 0; }

}  // namespace

void parse_argv_host(prism_context* c, int argc, char** argv) {
    if (argc == 1) usage();
    auto has = [&](int i, int k) { return i + k < argc; };
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (is(a, "-h")) {
            usage();
            std::exit(0);
        } else if (is(a, "-f")) {
            c->dtype = F4;
        } else if (is(a, "-d")) {
            c->dtype = F8;
        } else if (is(a, "-A")) {
            c->error_mode = ABS;
            if (has(i, 1)) c->eb = std::stod(argv[++i]);
        } else if (is(a, "-R")) {
            c->error_mode = REL;
            if (has(i, 1)) c->rel_eb = std::stod(argv[++i]);
        } else if (is(a, "-PW_R")) {  // parsed, not used by the method
            c->error_mode = PW_REL;
            if (has(i, 1)) c->pw_rel_eb = std::stod(argv[++i]);
        } else if (is(a, "-z")) {
            c->isComp = true;
            if (has(i, 1) && argv[i + 1][0] != '-') c->cmpFilePath = argv[++i];
        } else if (is(a, "-x")) {
            c->isDecomp = true;
            if (has(i, 1) && argv[i + 1][0] != '-') c->decFilePath = argv[++i];
        } else if (is(a, "-i")) {
            if (has(i, 1)) c->oriFilePath = argv[++i];
        } else if (is(a, "-1") || is(a, "-2") || is(a, "-3")) {
            const int nd = a[1] - '0';
            c->ndim = nd;
            if (!has(i, nd)) {
                usage();
                continue;
            }
            uint32_t* dims[3] = {&c->x, &c->y, &c->z};
            for (int k = 0; k < nd; ++k) *dims[k] = (uint32_t)std::stoi(argv[i + 1 + k]);
            i += nd;
        } else if (is(a, "-errors")) {
            c->compMode = 1;
            if (!has(i, 1)) throw std::invalid_argument("Number of error bounds not specified.");
            int k = 0;
            try {
                k = std::stoi(argv[++i] + 1);  // "-<nums>"
            } catch (const std::exception& e) {
                std::cerr << e.what() << '\n';
            }
            c->target_ebs.assign(k > 0 ? k : 0, 0.0);
            for (int j = 0; j < k; ++j) {
                if (!has(i, 1)) throw std::invalid_argument("Not enough error bounds provided.");
                c->target_ebs[j] = std::stod(argv[++i]);
            }
        } else if (is(a, "--report")) {
            if (!has(i, 1)) {
                std::cerr << "Error: missing argument after --report\n";
                return;
            }
            std::stringstream ss(argv[++i]);
            std::string tok;
            while (std::getline(ss, tok, ',')) {
                if (tok == "time") c->report_time = true;
                else if (tok == "cr") c->report_cr = true;
                else std::cerr << "Warning: unknown report option: " << tok << std::endl;
            }
        } else if (is(a, "-prog")) {
            c->bt = NB;
        } else if (is(a, "-cubic")) {  // test flag; not stored in the stream
            if (has(i, 1) && is(argv[++i], "natural"))
                for (bool& v : c->intp_param.use_natural) v = true;
        } else if (is(a, "-md")) {  // test flag; the auto-tuner decides
            if (has(i, 1)) {
                if (is(argv[i + 1], "true")) for (bool& v : c->intp_param.use_md) v = true;
                else if (is(argv[i + 1], "false")) for (bool& v : c->intp_param.use_md) v = false;
                ++i;
            }
        } else if (is(a, "-reverse")) {
            for (bool& v : c->intp_param.reverse) v = true;
        } else {
            std::printf("Error: Unrecognized option '%s'.\n", a);
            std::exit(EXIT_FAILURE);
        }
    }
    c->size = c->x * c->y * c->z * c->w;
    check(c);
}

void parse_argv(prism_context* c, int argc, char** argv) {
    parse_argv_host(c, argc, argv);
    if (hipMalloc((void**)&c->begin, 4 * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&c->end, 4 * sizeof(int)) != hipSuccess ||
        hipMemset(c->begin, 0, 4 * sizeof(int)) != hipSuccess || hipMemset(c->end, 0, 4 * sizeof(int)) != hipSuccess) {
        std::cerr << "[HIP ERROR] cannot allocate the retrieval state" << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

void release_context(prism_context* c) {
    if (c->begin) (void)hipFree(c->begin);
    if (c->end) (void)hipFree(c->end);
    c->begin = c->end = nullptr;
}
