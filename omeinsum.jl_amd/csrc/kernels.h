// Host-side launchers of the device kernels (all enqueue on the given stream).
#pragma once
#include "common.h"

namespace ome {

// Dry-run mode (thread-local): launchers count the kernels they would enqueue and return
// without launching.  Used to size a plan's scratch memory before it is allocated.
struct LaunchMode {
  bool dry = false;
  int64_t count = 0;
};
LaunchMode& launch_mode();
#define OME_DRY_RETURN(n)            \
  if (launch_mode().dry) {           \
    launch_mode().count += (n);      \
    return;                          \
  }

// Kernel timing (process-wide, off by default).  While on, a TimedScope brackets the kernels
// of one class with HIP events on their launch stream; scopes on a stream that is being
// captured into a graph record nothing.
enum KClass : int {
  KC_GEMM_PACKED = 0, KC_PACK = 1, KC_GEMM_MFMA = 2, KC_GEMM_VALU = 3,
  KC_PERMUTE = 4, KC_WALK = 5, KC_GROUPED = 6, KC_OTHER = 7, KC_COUNT = 8
};
void timing_enable(bool on);
bool timing_enabled();
// Waits for the recorded events of class kc, returns their summed duration and count, and
// clears them.
void timing_collect(int kc, double* total_ms, int64_t* launches);
struct TimedScope {
  int kc;
  hipStream_t st;
  hipEvent_t start = nullptr;
  TimedScope(int kc, hipStream_t st);
  ~TimedScope();
  TimedScope(const TimedScope&) = delete;
  TimedScope& operator=(const TimedScope&) = delete;
};

enum Prec : int { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP64 = 2 };
enum GemmKernel : int { GEMM_AUTO = -1, GEMM_VALU_TILE = 0, GEMM_VALU_DIRECT = 1, GEMM_MFMA = 2 };

void launch_axpby(int dt, const void* x, void* y, int64_t n, const Scal& sx, const Scal& sy,
                  int conj, hipStream_t st);
void launch_fill(int dt, void* y, int64_t n, const Scal& v, hipStream_t st);
void launch_acc_f64(int dt, const void* x, double* acc, int64_t n, hipStream_t st);
void launch_from_f64(int dt, const double* acc, void* y, int64_t n, const Scal& sx,
                     const Scal& sy, hipStream_t st);
// x <- x / max|x|; logacc[0] += log(max|x|).  work needs 1 + 1024 doubles.
void launch_normalize(int dt, void* x, int64_t n, double* work, double* logacc, hipStream_t st);
void launch_permute(int dt, const PermDesc& d, hipStream_t st);
void launch_walk(int dt, const WalkDesc& d, hipStream_t st);

int64_t gemm_valu_workspace(int dt, const GemmDesc& d, int ksplit);
void launch_gemm_valu(int dt, GemmDesc d, int kernel, int ksplit, hipStream_t st);

// Few outputs (M*N*batch <= 1024), long K: one workgroup per output and K part.
int dot_plan_parts(const GemmDesc& d);
void launch_gemm_dot(int dt, GemmDesc d, int parts, hipStream_t st);

// GEMMs on the real matrix cores in another arithmetic than the element type: complex
// products as real GEMMs (4M realified, 3M Gauss) and Float32 widened to Float64.
// prep_a / prep_b write the real operands (sizes from prep_bytes), the real GEMM runs
// on them, combine applies y <- sy*y + sx*C from the real result.
enum RealAlgo : int { RA_WIDEN = 0, RA_4M = 1, RA_3M = 2 };
struct RealShape {
  int64_t M, N, K, B;  // of the real GEMM
  int64_t a_elems, b_elems, c_elems;
};
RealShape real_shape(int algo, const GemmDesc& d);
void launch_prep_a(int dt, bool rdouble, int algo, const GemmDesc& d, void* out, hipStream_t st);
void launch_prep_b(int dt, bool rdouble, int algo, const GemmDesc& d, void* out, hipStream_t st);
void launch_combine(int dt, bool rdouble, int algo, const GemmDesc& d, const void* c,
                    hipStream_t st);

int mfma_plan_ksplit(const GemmDesc& d, int requested);
int64_t mfma_workspace(int dt, const GemmDesc& d, int ksplit);
void launch_gemm_mfma(int prec, GemmDesc d, int ksplit, hipStream_t st);

// Packed BF16x3 GEMM (Float32, or ComplexF32 realified into one real GEMM).
int64_t packed_bytes_a(int64_t M, int64_t N, int64_t K, int64_t batch);
int64_t packed_bytes_b(int64_t M, int64_t N, int64_t K, int64_t batch);
int64_t packed_tiles_mn(int64_t M, int64_t N);
bool packed_plain_overwrite(const GemmDesc& g);
void launch_gemm_packed(int dt, const GemmDesc& g, void* pa, void* pb, float* work, int ksplit,
                        hipStream_t st);

}  // namespace ome
