// qsim_fft_plan.h -- host-side description of one FFT axis pass, shared by
// the standalone plans and the fused solver jobs.
#pragma once

#include "qsim_fft_kernels.h"

namespace qsim {

struct FftAxis {
    size_t n = 1;       // line length
    size_t S = 1;       // stride between consecutive line elements
    size_t nlines = 1;  // number of lines
    bool p2 = false;    // register-resident power-of-two kernel
    int L = 1;          // generic kernel: lines per workgroup
    size_t lds = 0;     // generic kernel: dynamic LDS bytes
    GenFactors gf{};
    const void *tw = nullptr;  // twiddle table exp(-2 pi i m / n), m < n
    // four-step form of a long innermost axis, n = n1 * n2 (two passes
    // through a scratch tensor of the whole size)
    bool four = false;
    size_t n1 = 0, n2 = 0;
    const void *tw1 = nullptr, *tw2 = nullptr;
    void *tmp = nullptr;  // owned by the plan that prepared the axis
};

int twiddle_table(size_t n, int precision, const void **out);
bool is_p2_supported(size_t n, int precision);
int prepare_axis(FftAxis *a, size_t n, size_t S, size_t nlines, int precision);
// free what prepare_axis allocated (the four-step scratch)
void release_axis(FftAxis *a);
// Transform every line of the axis, in -> out, multiplying by `scale`.
int run_axis(const FftAxis &a, int precision, bool inv, const void *in, void *out, double scale,
             hipStream_t s);

}  // namespace qsim
