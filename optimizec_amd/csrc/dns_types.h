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

// FP32 complex values (8 bytes, as cf) whose line transforms run every radix stage of
// radix <= 16 as a batched DFT-matrix product on the matrix cores with FP32 accumulation:
//   cm3 -- three BF16 products per stage (F_hi x_hi + F_hi x_lo + F_lo x_hi)
//   cmh -- one FP16 product per stage, each butterfly's inputs scaled by a power of two
struct cm3 {
    float x, y;
};
struct cmh {
    float x, y;
};

// FP32 complex values of vector-stage transforms (as cf) whose intermediates between the
// two kernels of a step, V and G, are stored as FP16 pairs under a power-of-two scale per
// stored line (4 bytes per value; see the fused kernels)
struct cs {
    float x, y;
};

}  // namespace dnsfft
