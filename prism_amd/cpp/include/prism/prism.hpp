// prism.hpp -- entry point of the PRISM driver (the reference's
// include/prism.hpp): runs the compression and/or (progressive)
// decompression that a parsed prism_context describes.
#pragma once

#include "parameters.hpp"

void initialize(prism_context* config);
