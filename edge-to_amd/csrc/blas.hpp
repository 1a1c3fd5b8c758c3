// Deterministic reductions shared by the PCG loop, the OC update and the C ABI.
#pragma once
#include "common.hpp"

namespace et {

constexpr int kRedBlocks = 1024;  // fixed first-stage grid of every reduction

// partial: kRedBlocks doubles of device scratch; out: device scalar
int dot(int64_t n, const double* a, const double* b, double* partial, double* out, hipStream_t s);
int reduce_finalize(const double* partial, double* out, hipStream_t s);
// first stage only: kRedBlocks partials of a.b
int dot_partials(int64_t n, const double* a, const double* b, double* partial, hipStream_t s);
// the same over the first m partials
int reduce_finalize_n(const double* partial, int m, double* out, hipStream_t s);

}  // namespace et
