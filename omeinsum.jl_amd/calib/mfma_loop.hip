// Sustained BF16 MFMA rate of the packed BF16x3 GEMM's inner loop, without global memory.
//
// The loop is the one of k_gemm_packed_bf16x3_pipe<256, 256, 2, 4>: eight waves of 128 x 64,
// two LDS stages of 64 KiB, per 16-deep half-step 12 ds_read_b128 fragment reads and 24
// v_mfma_f32_32x32x16_bf16, fragments double-buffered, one workgroup barrier per 32-deep step,
// 256 workgroups (one per CU).  Only the global -> LDS copies and the epilogue are left out.
//
// The LDS stages are filled once, either with a constant pattern or with random bf16 values
// (random sign and mantissa, exponent in [2^-4, 2^4)).  The matrix cores draw more power on
// random operands, and the part lowers its clock under that load, so the two rates bound what
// any kernel with this loop can reach on this part with real data.
//
// Output: one JSON line {"steps": S, "const_bf16_tflops": C, "random_bf16_tflops": R,
// "random_bf16x3_tflops": R / 3}.  Each figure is the best of five launches after a warm-up.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int TM = 256, TN = 256, WM = 2, WN = 4;
constexpr int MB = TM / WM / 32, NB = TN / WN / 32;
constexpr int A_BYTES = TM * 32 * 2 * 2, B_BYTES = TN * 32 * 2 * 2, STAGE = A_BYTES + B_BYTES;
constexpr int GRID = 256, BLOCK = 64 * WM * WN;

template <bool RANDOM>
__global__ __launch_bounds__(BLOCK) void k_mfma_loop(float* out, int steps) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * STAGE];
  const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
  for (int i = t; i < 2 * STAGE / 4; i += BLOCK) {
    uint32_t v;
    if (RANDOM) {
      uint32_t x = (uint32_t)i * 2654435761u + blockIdx.x * 97u;
      x ^= x >> 13;
      x *= 0x5bd1e995u;
      x ^= x >> 15;
      const uint32_t y = x * 0x9E3779B1u;
      const uint32_t lo = (x & 0x807Fu) | ((123u + ((x >> 7) & 7u)) << 7);
      const uint32_t hi = ((y >> 16) & 0x807Fu) | ((123u + ((y >> 23) & 7u)) << 7);
      v = lo | (hi << 16);
    } else {
      v = 0x3F803F80u;  // two bf16 ones
    }
    ((uint32_t*)lds)[i] = v;
  }
  __syncthreads();
  const int wm = (w % WM) * (TM / WM), wn = (w / WM) * (TN / WN), jl = lane >> 5;
  // the packed kernel's fragment addresses (rows of 64 bytes, 16-byte chunks XOR-swizzled)
  auto rd = [&](int sb, int h, bf16x8 (&ah)[NB], bf16x8 (&al)[NB], bf16x8 (&bh)[MB],
                bf16x8 (&bl)[MB]) __attribute__((always_inline)) {
    const int base = sb * STAGE;
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
      const int r = wn + nb * 32 + (lane & 31);
      const int off = base + A_BYTES + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
      ah[nb] = *(const bf16x8*)(lds + off);
      al[nb] = *(const bf16x8*)(lds + off + TN * 64);
    }
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
      const int r = wm + mb * 32 + (lane & 31);
      const int off = base + r * 64 + 16 * ((2 * h + jl) ^ ((r >> 2) & 3));
      bh[mb] = *(const bf16x8*)(lds + off);
      bl[mb] = *(const bf16x8*)(lds + off + TM * 64);
    }
  };
  f32x16 acc[NB][MB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  auto mm = [&](const bf16x8 (&ah)[NB], const bf16x8 (&al)[NB], const bf16x8 (&bh)[MB],
                const bf16x8 (&bl)[MB]) __attribute__((always_inline)) {
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[nb], bh[mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bl[mb], acc[nb][mb], 0, 0, 0);
#pragma unroll
    for (int nb = 0; nb < NB; ++nb)
#pragma unroll
      for (int mb = 0; mb < MB; ++mb)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[nb], bh[mb], acc[nb][mb], 0, 0, 0);
  };
  bf16x8 ah0[NB], al0[NB], bh0[MB], bl0[MB], ah1[NB], al1[NB], bh1[MB], bl1[MB];
  rd(0, 0, ah0, al0, bh0, bl0);
  for (int kt = 0; kt < steps; ++kt) {
    const int sb = kt & 1;
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __builtin_amdgcn_sched_barrier(0);
    rd(sb, 1, ah1, al1, bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
    mm(ah0, al0, bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_sched_barrier(0);
    rd(sb ^ 1, 0, ah0, al0, bh0, bl0);
    __builtin_amdgcn_sched_barrier(0);
    mm(ah1, al1, bh1, bl1);
    __builtin_amdgcn_sched_barrier(0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) s += acc[i][j][r];
  out[blockIdx.x * BLOCK + t] = s;
}

template <bool RANDOM>
double best_tflops(float* out, int steps) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.0;
  double best = 0.0;
  for (int rep = 0; rep < 6; ++rep) {  // the first launch warms the part up and is not counted
    if (hipEventRecord(e0) != hipSuccess) return -1.0;
    hipLaunchKernelGGL((k_mfma_loop<RANDOM>), dim3(GRID), dim3(BLOCK), 0, 0, out, steps);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
    // 48 MFMAs of 2 * 32 * 32 * 16 flop per wave per step
    const double flop = 48.0 * 32768.0 * steps * (BLOCK / 64) * GRID;
    if (rep > 0 && flop / ms / 1e9 > best) best = flop / ms / 1e9;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

int main(int argc, char** argv) {
  int steps = 512;
  if (argc > 1 && sscanf(argv[1], "%d", &steps) != 1) steps = 512;
  if (steps < 1 || steps > (1 << 20)) {
    fprintf(stderr, "steps must lie in [1, 2^20]\n");
    return 2;
  }
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * GRID * BLOCK) != hipSuccess) {
    fprintf(stderr, "hipMalloc failed\n");
    return 1;
  }
  const double c = best_tflops<false>(out, steps);
  const double r = best_tflops<true>(out, steps);
  (void)hipFree(out);
  if (c < 0 || r < 0) {
    fprintf(stderr, "HIP call failed\n");
    return 1;
  }
  printf("{\"steps\": %d, \"const_bf16_tflops\": %.1f, \"random_bf16_tflops\": %.1f, "
         "\"random_bf16x3_tflops\": %.1f}\n", steps, c, r, r / 3.0);
  return 0;
}
