// Part 2 of the longitude FFT: the FP32-arithmetic entry points on FP32 grids
// (the kernels and their description are in zgl_fft.hip).
#define ZGL_FFT_PART 2
#include "zgl_fft.hip"
