// dns_types.h -- plain data types shared by host and device code of the library.
#pragma once

namespace dnsfft {

// FP64 complex value as stored in HBM and LDS (16 bytes, re then im)
struct cd {
    double x, y;
};

// FP32 complex value (8 bytes) of the mixed-precision transforms
struct cf {
    float x, y;
};

}  // namespace dnsfft
