// band.hpp -- two-stage reduction of the batched FP32 eigensolver
// (k_band.hip): dense -> band -> tridiagonal and the back-transform of the
// bulge-chasing reflectors.
#pragma once

#include "common.hpp"

namespace cir {

constexpr int kBandNB = 32;           // band half-width (panel width of stage 1)
constexpr int kBandW = 2 * kBandNB;   // stored band row: r - c in [0, 2 NB)
constexpr int kQ2G = 16;              // sweeps per compact-WY block of the bulge-chasing reflectors
constexpr int kQ2R = 48;              // rows of such a block (NB + kQ2G - 1, padded)
constexpr int kQ2S = 16;              // eigenvector columns per workgroup in the Q2 back-transform

// Largest n of the stage-1 register panel (4 rows of 512 threads below the band).
inline int sb_max_n() { return 4 * 512 + kBandNB; }

// Tasks (block positions) of bulge-chasing sweep s: rows s+1+k NB, k < count.
__host__ __device__ inline int sb_tasks(int n, int s) { return (s <= n - 3) ? (n - 2 - s) / kBandNB + 1 : 0; }
// Compact-WY blocks of the bulge-chasing reflectors: block (g, k) holds the
// reflectors of sweeps g kQ2G .. g kQ2G + kQ2G - 1 at position k; its index is
// q2_block_offset(n, g) + k, and group g has sb_tasks(n, g kQ2G) blocks.
__host__ __device__ inline int q2_groups(int n) { return n >= 3 ? (n - 2 + kQ2G - 1) / kQ2G : 0; }
__host__ __device__ inline int64_t q2_block_offset(int n, int g) {
    int64_t o = 0;
    for (int q = 0; q < g; ++q) o += sb_tasks(n, q * kQ2G);
    return o;
}
inline int64_t q2_blocks(int n) { return q2_block_offset(n, q2_groups(n)); }
// floats per block: V (kQ2R x kQ2G, column per reflector), tau (kQ2G), -V T (kQ2R x kQ2G)
constexpr int kQ2BlockFloats = 2 * kQ2R * kQ2G + kQ2G;

// Stage 1 for a batch: A (n x n, ld, lower triangle used) -> band rows BND
// (n x kBandW each), reflectors in A and tau (n each).  Work arrays per
// matrix: P n x 2NB, Yh n x NB, Mm NB x NB.
int sy2sb(int batch, int n, int ld, float* A, float* BND, float* tau, float* P, float* Yh, float* Mm, hipStream_t s);
// Stage 2: band rows -> d, e (n each; e[n-1] = 0) and the bulge-chasing
// reflectors by blocks Q2 (q2_blocks(n) * kQ2BlockFloats).
int sb2st(int batch, int n, float* BND, float* d, float* e, float* Q2, hipStream_t s);
// The blocks of Q2 into compact-WY form (T factors, operand order of
// q2_apply), in place.
int q2_prep(int batch, int n, float* Q2, hipStream_t s, bool bf16);
// Z <- Q2 Z for the eigenvectors Z of the tridiagonal matrix (row-major,
// Z[i * ld + k] = component i of vector k), batch stride n * ld; Q2 as
// q2_prep leaves it.
// bf16: the blocks' products as split-BF16x3 MFMA (q2_prep with the same flag).
int q2_apply(int batch, int n, int ld, const float* Q2, float* Z, hipStream_t s, bool bf16);

}  // namespace cir
