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

// roctx ranges (process-wide, off by default), resolved from the ROCm profiler's roctx library
// when switched on.  RangeScope pushes a nested range on the calling thread; range_start /
// range_stop bracket work that ends on another call (a request from admission to retirement).
bool ranges_enable(bool on);  // false when the roctx library cannot be loaded
bool ranges_enabled();
uint64_t range_start(const char* name);  // 0 when ranges are off
void range_stop(uint64_t id);
struct RangeScope {
  bool on;
  explicit RangeScope(const char* name);
  ~RangeScope();
  RangeScope(const RangeScope&) = delete;
  RangeScope& operator=(const RangeScope&) = delete;
};

// Writes code into *status (a device-visible word) from a kernel on st.
void launch_status_set(int* status, int code, hipStream_t st);

enum Prec : int { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_FP64 = 2, PREC_BF16X6 = 3 };
enum GemmKernel : int { GEMM_AUTO = -1, GEMM_VALU_TILE = 0, GEMM_VALU_DIRECT = 1, GEMM_MFMA = 2 };

void launch_axpby(int dt, const void* x, void* y, int64_t n, const Scal& sx, const Scal& sy,
                  int conj, hipStream_t st);
void launch_fill(int dt, void* y, int64_t n, const Scal& v, hipStream_t st);
void launch_acc_f64(int dt, const void* x, double* acc, int64_t n, hipStream_t st);
void launch_from_f64(int dt, const double* acc, void* y, int64_t n, const Scal& sx,
                     const Scal& sy, hipStream_t st);
// x <- x / max|x|; logacc[0] += log(max|x|).  work needs 1 + 1024 doubles.
void launch_normalize(int dt, void* x, int64_t n, double* work, double* logacc, hipStream_t st);
// Float32 / ComplexF32 x -> Float64 / ComplexF64 y = x / max|x|; log(max|x|) is added to
// logacc[0] (work: the normalisation scratch of launch_normalize)
void launch_widen_normalized(int dt, const void* x, void* y, int64_t n, double* work,
                             double* logacc, hipStream_t st);
void launch_permute(int dt, const PermDesc& d, hipStream_t st);
// Two independent permutes (a GEMM's two operands); one launch when both take the
// bit-permutation path, else one after the other.
void launch_permute2(int dt, const PermDesc& d0, const PermDesc& d1, hipStream_t st);
void launch_walk(int dt, const WalkDesc& d, hipStream_t st);

int64_t gemm_valu_workspace(int dt, const GemmDesc& d, int ksplit);
void launch_gemm_valu(int dt, GemmDesc d, int kernel, int ksplit, hipStream_t st);

// Binary contraction with a short contracted space read in place through per-label strides:
// y[o] = sum_k A[fa(o) + ka(k)] * B[fb(o) + kb(k)], one output per thread, the contracted
// offsets tabulated in LDS per block.  Replaces permute + GEMM (+ permute) for the many small
// steps of tensor-network trees, whose labels are scattered over the operands.
constexpr int SG_MAXO = 40;     // output labels
constexpr int SG_MAXK = 16;     // contracted labels
constexpr int SG_KCAP = 1024;   // contracted points
struct StridedGemmDesc {
  const void* a;
  const void* b;
  void* y;
  int64_t nout;
  int32_t nk, no, ni;
  FDiv odiv[SG_MAXO];  // output labels in y's order (first fastest)
  int32_t oa[SG_MAXO], ob[SG_MAXO], oy[SG_MAXO];
  FDiv kdiv[SG_MAXK];  // contracted labels (first fastest)
  int32_t ka[SG_MAXK], kb[SG_MAXK];
  Scal sx, sy;
  int mode;
};
void launch_gemm_strided(int dt, const StridedGemmDesc& d, hipStream_t st);

// Small tree steps fused into one launch: workgroup b evaluates subtree b, each step the
// strided contraction above with its output overwritten.  A subtree runs as a sequence of
// chunks, a workgroup barrier after each: a chunk is up to SF_LMAX steps of one level (a step's
// level is one more than the highest level among its fused operands; steps whose operands are
// all leaves have level 1), so its steps read only leaves and outputs of earlier chunks.
// Operands and outputs live in the workgroup's LDS arena (SF_LDS bytes) where they fit: the
// outputs inside the subtree, and copies of the leaves read above level 1, made by copy steps
// (y = the leaf) at level 1; the rest are leaves (pointers passed with the launch) or blocks of
// the plan's arena, as is every subtree root's output.
// A chunk is a range of work items, each step's from a wave boundary, 2^lg items per output
// (they split its contracted points).  Item i of a step works on output o = i >> lg (written at
// element o of y: outputs are enumerated in y's order), whose operand element offsets come from
// the chunk's item table (a 32-bit word per item: A's offset in the low half, B's in the high).
// steps holds every subtree's descriptors in chunk order.  first is one int32 array: nsub + 1
// entries, subtree q's chunks being [first[q], first[q + 1]); then, for nch = first[nsub]
// chunks, nch + 1 chunk starts (chunk c is descriptors [e[c], e[c + 1])); then nch + 1 item
// table starts (chunk c's table is the words [f[c], f[c + 1]) of the same array); then the
// item tables.
constexpr int SF_MAXNK = 64;     // contracted points
constexpr int SF_MAXOUT = 8192;  // output points
constexpr int SF_MAXLEAF = 256;  // leaves of a tree whose steps may be fused
constexpr int SF_THREADS = 1024; // threads of the workgroup that runs one subtree
constexpr int SF_LMAX = 16;      // steps of one chunk
constexpr int SF_LDS = 96 * 1024;  // LDS arena of a subtree, bytes
enum : int32_t { FS_A_LDS = 1, FS_B_LDS = 2, FS_Y_LDS = 4, FS_COPY = 8 };
struct alignas(16) FusedStep {
  int32_t a_leaf, b_leaf;        // >= 0: leaf index, read in place; -1: in the arena or LDS
  int32_t flags;                 // FS_*: operands / output in the LDS arena; copy step (y = a)
  int32_t nout, nk;
  int32_t ibase;                 // first work item of the step in its chunk (a wave boundary)
  int32_t total, lg;             // the chunk's work items, and log2 of the items per output
  int64_t a_off, b_off, y_off;   // byte offsets in the plan's arena, or in the LDS arena
  int64_t pad_;
  uint16_t ka[SF_MAXNK], kb[SF_MAXNK];  // operand element offsets of each contracted point
};
static_assert(sizeof(FusedStep) % 16 == 0, "16-byte descriptors");
struct FusedLeaves {
  const void* p[SF_MAXLEAF];
};
void launch_fused_steps(int dt, const FusedStep* steps, const int32_t* first, int nsub,
                        const FusedLeaves& leaves, char* arena, hipStream_t st);
// The fused subtrees of several plans (several requests of the request loop) in one launch:
// job j owns workgroups [sub_begin, sub_begin + nsub) and evaluates its plan's subtrees (steps
// and first as above) on its own arena and leaves (a device table of leaf pointers).  jobs is
// a device array.
struct FusedJob {
  const FusedStep* steps;
  const int32_t* first;
  char* arena;
  const void* const* leaves;
  int32_t sub_begin, nsub;
};
constexpr int SF_MAXJOBS = 64;
void launch_fused_jobs(int dt, const FusedJob* jobs, int njobs, int nsub_total, hipStream_t st);

// Few outputs (1024 < M*N*batch <= 16384), K >= 512, at most 2^26 products, not complex: one
// launch, a workgroup per 4x4 (or 1x16, 16x1) output tile over all of K, reduced in a fixed
// order.
bool panel_fits(int dt, const GemmDesc& d);
void launch_gemm_panel(int dt, GemmDesc d, hipStream_t st);

// Few outputs (M*N*batch <= 1024), long K: one workgroup per output and K part.
int dot_plan_parts(const GemmDesc& d);
int64_t dot_workspace(int dt, const GemmDesc& d, int parts);
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

// Grouped launches (one persistent launch over many problems).
struct GGProb {  // one problem of a grouped GEMM: dense column-major, batch-major
  const void* a;
  const void* b;
  void* c;
  int64_t m, n, k;
  int32_t tiles_m, tiles_n;
  int32_t vec_a, vec_b;  // 16-byte vector loads allowed along m (A) / k (B)
  FDiv per_div, tm_div;  // invariant division by tiles_m * tiles_n and by tiles_m
  int64_t first_tile;    // the problem's first tile in the launch's tile numbering
};
// Workgroup b runs on XCD b % GG_XCDS (the dispatcher deals workgroups to the XCDs in turn);
// the tile list is cut into one contiguous range per XCD, each holding whole problems, so a
// problem's operand panels are fetched into one XCD's L2.  Launch grids are multiples of
// GG_XCDS.
constexpr int GG_XCDS = 8;
struct GGDesc {
  const GGProb* probs;   // device
  const int64_t* first;  // device: first tile of each problem (ascending)
  const int32_t* map;    // device: problem of each tile
  int nprob;
  int64_t total_tiles;
  Scal alpha, beta;
  int mode;
  int64_t xcd_first[GG_XCDS + 1];  // tiles of XCD c: [xcd_first[c], xcd_first[c + 1])
  int offsets32;  // every problem has M, N, K < 2^22 and batch slices of A, B, C < 2^31 bytes
};
enum GGVariant : int {
  GG_F32_64 = 0, GG_F32_64_BK32 = 1, GG_F32_128 = 2, GG_BF16X3_64 = 3, GG_BF16X3_128 = 4,
  GG_F32_W32 = 5, GG_BF16X3_W32 = 6, GG_BF16X3_64_BK64 = 7, GG_F32_RUN = 8, GG_BF16X3_RUN = 9,
  GG_BF16X6_64 = 10,
  GG_NVARIANTS = 11
};
// Workgroups per CU of the run variants (persistent, one per SIMD wave slot they fit: 20 KiB
// of LDS each; BF16x3 in 72 VGPRs, FP32 in 80).
constexpr int gg_run_per_cu(bool x3) { return x3 ? 7 : 6; }
void gg_tile_shape(int dt, int variant, int& tm, int& tn);
void launch_grouped_gemm(int dt, int variant, const GGDesc& g, int grid, hipStream_t st);

constexpr int GP_MAXR = 24;
constexpr int GP_MAXG = 8;  // dims in a permute tile's input or output group
constexpr int GP_SCAP_BYTES = 32768;  // LDS bytes of a shared-dim permute tile
// One permute.  Plain path (tiled == 0): output dims (fastest first) with their input strides.
// Tiled path: the leading output dims (group O, linear index a, contiguous in y) and the
// leading input dims (group I, linear index b, contiguous in x) are disjoint; a block moves
// a tla x tlb (a, b) tile through LDS for one value of the remaining dims (group R), so both
// the reads (along b) and the writes (along a) are coalesced.
struct GPProb {
  const void* x;
  void* y;
  int64_t n;
  int32_t rank;
  int32_t tiled;
  FDiv odiv[GP_MAXR];        // plain: output extents
  int64_t istride[GP_MAXR];  // plain: input stride per output dim
  int32_t no, ni, nr;        // tiled: dims in O, I, R
  int64_t pa, pb;            // tiled: extent of a (product over O) and of b (over I)
  int64_t ta, tb;            // tiled: tiles along a and b
  int32_t tla, tlb;          // tiled: tile extent along a and b (tla * tlb <= 4096)
  FDiv tlb_div, tla_div;
  FDiv odv[GP_MAXG];         // O dims (output order), their input strides
  int64_t o_is[GP_MAXG];
  FDiv idv[GP_MAXG];         // I dims (input order), their output strides
  int64_t i_os[GP_MAXG];
  FDiv rdv[GP_MAXR];         // R dims, input and output strides
  int64_t r_is[GP_MAXR], r_os[GP_MAXR];
  // tiled == 2: the whole of I and O and a chunk of tsc values of the shared dim S, the dim
  // that follows I in the input and O in the output (runs of pb * tsc and pa * tsc elements)
  int32_t tsc;
  FDiv tsc_div;
  FDiv st_div;  // tiled == 2: division by tsc * pb, the input run of one a
  int64_t s_ext, s_is, s_os, ts_tiles;
};
struct GPDesc {
  const GPProb* probs;
  const int64_t* first;  // first block of each problem
  const int32_t* map;    // problem of each block
  int nprob;
  int64_t total_blocks;
  Scal sx, sy;
  int mode;
  int64_t blk_begin, blk_end;  // the block range of one launch
};
// blocks [0, split) are the shared-dim problems (tiled == 2), [split, total) the others
void launch_grouped_permute(int dt, const GPDesc& g, int64_t split, hipStream_t st);

int mfma_plan_ksplit(const GemmDesc& d, int requested);
int64_t mfma_workspace(int dt, const GemmDesc& d, int ksplit);
void launch_gemm_mfma(int prec, GemmDesc d, int ksplit, hipStream_t st);
// ComplexF32 at FP32 on the matrix cores, realified (4M) while staged; A needs a unit stride
// along m or k and B along k or n (complex elements).
int mfma_c32_plan_ksplit(const GemmDesc& d);
int64_t mfma_c32_workspace(const GemmDesc& d, int ksplit);
void launch_gemm_mfma_c32(GemmDesc d, int ksplit, hipStream_t st);

// Packed split-precision GEMM (Float32, or ComplexF32 realified into one real GEMM) on np
// bf16 planes per operand: np = 2 (BF16x3, three products) or 3 (BF16x6, six products).
int64_t packed_bytes_a(int64_t M, int64_t N, int64_t K, int64_t batch, int np);
int64_t packed_bytes_b(int64_t M, int64_t N, int64_t K, int64_t batch, int np);
int64_t packed_tiles_mn(int64_t M, int64_t N, int np);
bool packed_plain_overwrite(const GemmDesc& g);
void launch_gemm_packed(int dt, const GemmDesc& g, void* pa, void* pb, float* work, int ksplit,
                        hipStream_t st, int np);

}  // namespace ome
