// Declarations of the parameter-file kernels (defined in params_io.h, compiled in
// zgr_params.hip) for the host code of zgr.hip, and the point sizes both sides use.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace zgr {

constexpr int G2_WORDS = 32;   // one affine G2 point in memory: x.c0 x.c1 y.c0 y.c1
constexpr int G2X_WORDS = 64;  // one XYZZ G2 point in the DFT workspace

__global__ void k_g2_fixed_base_table(uint32_t* table);
__global__ void k_power_tau(const uint32_t* tau_canon, uint64_t start, uint32_t n, const uint32_t* table1,
                            const uint32_t* table2, uint32_t* out1, uint32_t* out2, int* bad);
__global__ void k_vanish_closed(const uint32_t* pc, const uint32_t* consts, int log_n, int h, const uint32_t* table2,
                                uint32_t* out, int* bad);
__global__ void k_vanish_load(const uint32_t* g2pp, const uint32_t* w_mont, int scale, int log_n, int h,
                              uint32_t* work);
__global__ void k_vanish_stage(uint32_t* work, const uint32_t* tw_inv, int s, int log_n, int h);
__global__ void k_vanish_store(const uint32_t* work, int h, uint32_t* out, int* bad);
__global__ void k_g1_ark(const uint32_t* pts, uint32_t n, int mont, int compressed, uint32_t* out);
__global__ void k_g2_ark(const uint32_t* pts, uint32_t n, int mont, int compressed, uint32_t* out);
__global__ void k_fq_convert(const uint32_t* in, uint32_t* out, uint64_t count, int to_mont);

// ceremony import (ppot.h): decode and check count encoded points, writing canonical
// affine points and one status per point (0 = accepted)
constexpr int PPOT_G1_THREADS = 256;
constexpr int PPOT_G2_THREADS = 64;
__global__ void k_ppot_g1u(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status);
__global__ void k_ppot_g1c(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status);
__global__ void k_ppot_g2u(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status);
__global__ void k_ppot_g2c(const uint32_t* src, size_t count, uint32_t* out, uint32_t* status);

}  // namespace zgr
