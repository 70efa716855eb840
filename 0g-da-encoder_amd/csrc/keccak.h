// Keccak-f[1600] and the two fixed-length Keccak-256 sponges the Merkle tree needs
// (tiny-keccak 2.0.2 Keccak::v256: rate 136 bytes, original 0x01 padding).
// 64-bit lanes live in VGPR pairs; rotations compile to v_alignbit_b32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zgr {

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int s) {
    return (x << s) | (x >> (64 - s));
}

__device__ __forceinline__ void keccak_f1600(uint64_t A[25]) {
    constexpr uint64_t RC[24] = {
        0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
        0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
        0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
        0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
        0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
        0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
    // rho offsets and pi destinations along the (x, y) -> (y, 2x + 3y) walk from (1, 0)
    constexpr int ROT[24] = {1, 3, 6, 10, 15, 21, 28, 36, 45, 55, 2, 14,
                             27, 41, 56, 8, 25, 43, 62, 18, 39, 61, 20, 44};
    constexpr int PI[24] = {10, 7, 11, 17, 18, 3, 5, 16, 8, 21, 24, 4,
                            15, 23, 19, 13, 12, 2, 20, 14, 22, 9, 6, 1};
#pragma unroll 1
    for (int r = 0; r < 24; r++) {
        uint64_t C[5];
#pragma unroll
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            uint64_t D = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
#pragma unroll
            for (int y = 0; y < 25; y += 5) A[x + y] ^= D;
        }
        uint64_t cur = A[1];
#pragma unroll
        for (int t = 0; t < 24; t++) {
            uint64_t nxt = A[PI[t]];
            A[PI[t]] = rotl64(cur, ROT[t]);
            cur = nxt;
        }
#pragma unroll
        for (int y = 0; y < 25; y += 5) {
            uint64_t b0 = A[y], b1 = A[y + 1], b2 = A[y + 2], b3 = A[y + 3], b4 = A[y + 4];
            A[y] = b0 ^ (~b1 & b2);
            A[y + 1] = b1 ^ (~b2 & b3);
            A[y + 2] = b2 ^ (~b3 & b4);
            A[y + 3] = b3 ^ (~b4 & b0);
            A[y + 4] = b4 ^ (~b0 & b1);
        }
        A[0] ^= RC[r];
    }
}

// keccak256 of exactly 64 bytes given as 8 little-endian lanes
__device__ __forceinline__ void keccak256_64(const uint64_t in[8], uint64_t out[4]) {
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) A[i] = in[i];
    A[8] ^= 0x01ULL;
    A[16] ^= 0x8000000000000000ULL;
    keccak_f1600(A);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

// keccak256 of exactly 256 bytes given as 32 little-endian lanes (two blocks)
__device__ __forceinline__ void keccak256_256(const uint64_t in[32], uint64_t out[4]) {
    uint64_t A[25];
#pragma unroll
    for (int i = 0; i < 25; i++) A[i] = 0;
#pragma unroll
    for (int i = 0; i < 17; i++) A[i] = in[i];
    keccak_f1600(A);
#pragma unroll
    for (int i = 0; i < 15; i++) A[i] ^= in[17 + i];
    A[15] ^= 0x01ULL;
    A[16] ^= 0x8000000000000000ULL;
    keccak_f1600(A);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = A[i];
}

}  // namespace zgr
