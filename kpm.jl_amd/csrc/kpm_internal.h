// kpm_internal.h -- shared declarations of libkpm_hip.so (gfx950 only).
//
// Device data layout (HBM)
//   recurrence block  NH rows x NRp columns, ROW-interleaved: element (i, r)
//                     at i*NRp + r.  One row of a block is the NRp probe
//                     values of site i, so a CSR gather of neighbour j reads
//                     one contiguous NRp*elem-byte segment.  FP64 blocks hold
//                     double2 (re, im), FP32 blocks float2.  NRp >= NR is NR
//                     rounded up so a row is a power-of-two number (<= 64)
//                     of 16-byte words, or a multiple of 64 words; padded
//                     columns hold zeros and stay zero under the recurrence.
//   operator          CSR, int32 row pointers / column indices, values
//                     complex FP64 (double2) or real FP64 when Im == 0.
//   moments           NR x NC column-major double2 on the device, copied
//                     to the host once per engine call.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/kpm_hip.h"

namespace kpm {

struct PendingCombine;
struct Combine {
  int nd;
  int tcol[2];
  double alpha[2];
  int bcol[2];
  double bscale[2];
};

// ---- error plumbing -------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define KPM_HIP_TRY(expr)                                                          \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return ::kpm::fail(KPM_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define KPM_TRY(expr)        \
  do {                       \
    int _rc = (expr);        \
    if (_rc != KPM_OK) return _rc; \
  } while (0)

// ---- row geometry ----------------------------------------------------------
// A 16-byte word holds one complex double (FP64) or two complex floats (FP32).
inline int cplx_per_word(int prec) { return prec == KPM_VEC_F32 ? 2 : 1; }
inline int elem_bytes(int prec) { return prec == KPM_VEC_F32 ? 8 : 16; }
// words per row for NR probe columns
int row_words(int64_t NR, int prec);
inline int64_t padded_nr(int64_t NR, int prec) { return (int64_t)row_words(NR, prec) * cplx_per_word(prec); }

// ---- workspace cache ---------------------------------------------------------
struct Buffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace kpm

// State of the primitive recurrence path (kpm_cheb_step / kpm_spmm followed
// by columnwise_dot! calls): every primitive step also reduces <Vp|Vp> and
// <Vp|Vout> into per-workgroup partials (two halves, alternating), keyed on
// the step's blocks and the context's write epoch; a matching dot combines
// those partials instead of re-reading the blocks, and the combine is
// deferred into the next primitive step's launch when it can be.
struct kpm_prim_state {
  kpm::Buffer part;             // 2 halves of per-workgroup dot partials
  int k = 0;                    // primitive steps so far (selects the half)
  // dots of the last primitive step
  bool cached = false;
  const void* vp = nullptr;     // device data of Vp (d0 = <Vp|Vp>)
  const void* vout = nullptr;   // device data of Vout (d1 = <Vp|Vout>)
  uint64_t epoch = 0;
  void* part_ptr = nullptr;     // the half holding them
  int G = 0, NRp = 0, NR = 0;
  // combine waiting for the next primitive step (or a flush)
  bool pending = false;
  void* pend_mu = nullptr;      // accumulator data
  void* pend_part = nullptr;    // partials it reduces
  int pend_G = 0, pend_NRp = 0, pend_NR = 0;
  kpm::Combine pend_cb{2, {0, 0}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  bool pend_used[2] = {false, false};
};

// the forward pass kpm_1d_adjoint_forward keeps for kpm_1d_adjoint_reverse
struct kpm_adj1_state {
  bool valid = false;
  uint64_t H_serial = 0;
  int64_t NC = 0, NR = 0, NH = 0;
  int prec = 0;
  int64_t ckpt_req = 0;     // the caller's checkpoint stride
  int64_t stride_used = 0;  // 1: every vector kept; K > 1: checkpoint pairs
};

struct kpm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t events[16] = {};
  int num_cus = 0;
  // grow-only device scratch, keyed by role
  kpm::Buffer partials;   // per-block dot partial sums
  kpm::Buffer mu;         // device moment accumulator
  kpm::Buffer staging;    // host-layout block for upload/download
  kpm::Buffer gemm_part;  // split-K partials of the kpm_2d contraction
  kpm::Buffer work[4];    // engine-owned recurrence blocks
  kpm::Buffer lwork;      // bra block of the bra/ket engine
  kpm::Buffer slots_l;    // kpm_2d left slot array
  kpm::Buffer slots_r;    // kpm_2d right slot array
  kpm::Buffer ring[8];    // kpm_2d / kpm_3d ring and starter blocks (5-7: plane mode's left ring)
  kpm::Buffer pinned;     // page-locked host staging for operator uploads
  kpm::Buffer adjoint;    // adjoint workspace: v | beta ring | alpha store
  kpm::Buffer adj_grad;   // adjoint gradient accumulator (nnz complex)
  kpm::Buffer adj2;       // kpm_2d adjoint: chain slots, mixes, Clenshaw ring
  kpm::Buffer adj2_aux;   // kpm_2d adjoint: gradient accumulators, mix coefficients
  // merged operator of the dual pass: the union pattern of two operators
  // (rowptr | colind | first values | second values), rebuilt when the pair
  // of operator serials changes
  kpm::Buffer merged;
  uint64_t merged_key[2] = {0, 0};
  int merged_max_row = 0;
  int64_t merged_nnz = 0;
  int dual_pass = 1;      // fuse the kpm_2d / kpm_3d right chain with the Jb pass
  int bf16_planes = 1;    // split-BF16 mode: slots written as planes (kpm_2d, 128+ slots a side)
  int64_t merged_limit = -1;  // bytes the merged dual-pass operator may take (< 0: no limit)
  // launch / synchronisation counters (kpm_ctx_counters)
  int64_t counters[KPM_COUNT_N] = {};
  // incremented by every entry point that writes a caller's block
  uint64_t epoch = 1;
  // device address and size of the allocation-free path's scratch array
  // (noalloc.hip), looked up once per device at context creation
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  kpm_prim_state prim;
  kpm_adj1_state adj1;
};

// device moment accumulator (kpm_accum_*): NR x (NC + 1) double2,
// column-major; column NC is scratch for unused combine entries
struct kpm_accum {
  kpm_ctx* ctx = nullptr;
  int64_t NR = 0, NC = 0;
  void* d = nullptr;
};

// Entries past nnz that every CSR column-index and value array carries
// (zeros): a lane of the fused pass loads its share of a row's nonzeros with
// one vector load per array, which may read up to this far past the row's
// end (past the last row: past nnz).
constexpr int64_t kCsrPad = 32;

struct kpm_csr {
  kpm_ctx* ctx = nullptr;
  // position in this CSR of the k-th stored entry of the CSC arrays it was
  // built from (the adjoint returns gradients in the caller's CSC order)
  std::vector<int> csc_to_csr;
  int* csc_to_csr_dev = nullptr;  // device copy, uploaded by the first adjoint that needs it
  int64_t n = 0;
  int64_t nnz = 0;
  int is_real = 0;
  int max_row_nnz = 0;
  int sorted = 0;         // every row's column indices strictly increase
  // every row stores its diagonal among its first diag_slots nonzeros
  // (INT32_MAX when some row stores none)
  int diag_slots = INT32_MAX;
  uint64_t serial = 0;    // process-unique identity (never reused)
  int* rowptr = nullptr;  // n+1
  int* colind = nullptr;  // nnz
  void* vals = nullptr;   // nnz x (double2 | double)
  // the same values rounded to FP32, nnz x (float2 | float): what the
  // recurrence of FP32 blocks reads (attach_vals32, at creation)
  void* vals32 = nullptr;
};

struct kpm_block {
  kpm_ctx* ctx = nullptr;
  int64_t NH = 0;
  int64_t NR = 0;
  int64_t NRp = 0;
  int prec = KPM_VEC_F64;
  void* d = nullptr;
  bool owned = true;
};

namespace kpm {

int ensure(kpm_ctx* ctx, Buffer& b, size_t bytes);
// page-locked host staging (ctx->pinned, grow-only; synchronises the stream
// before it replaces a smaller buffer)
int ensure_pinned(kpm_ctx* ctx, size_t bytes);
// host CSR (0-based int64 rowptr/colind, interleaved complex values) -> device CSR
int upload_csr(kpm_ctx* ctx, int64_t n, const std::vector<int64_t>& rp, const std::vector<int64_t>& ci,
               const std::vector<double>& v_cplx, kpm_csr** out);
// CSC arrays -> device CSR through page-locked staging in one scatter pass
// (conj: build adjoint(A) instead of A); perm, when non-null, receives the
// CSR position of every CSC entry
int csr_from_csc_staged(kpm_ctx* ctx, int64_t n, const int64_t* colptr, const int64_t* rowval,
                        const double* nzval, int base, bool adjoint, std::vector<int>* perm,
                        kpm_csr** out);
// round A's values to FP32 into A->vals32 (device allocation + one
// conversion kernel on the context stream); every operator gets them at
// creation, so passes over FP32 blocks never allocate
int attach_vals32(kpm_ctx* ctx, kpm_csr* A);
// validate Julia CSC arrays (square n x n, index base 0 or 1)
int check_csc(int64_t n, const int64_t* colptr, const int64_t* rowval, int base);
// A non-owning block view over a cached workspace buffer.
int work_block(kpm_ctx* ctx, Buffer& buf, int64_t NH, int64_t NR, int prec, kpm_block* out);

// ---- kernels (spmm.hip) -------------------------------------------------------
// MODE_DBL_SEG: MODE_DBL on the segmented schedule of the batched engine
// MODE_SRC: out += gamma * l (no dots), the reverse Clenshaw step of the adjoint
// MODE_DUAL: also out2 = gamma * B * x from the same gathers, where A and B
// share one (merged) pattern and B's complex values are vals2 (no dots)
// MODE_DBLA: MODE_DBL with the dots summed over the probe columns as well
// (the probe-averaged moments of kpm_1d(...; avg_output = true)); one pair of
// partials per wave, [wave][d], so its partials count kBlock/64 per
// workgroup and row chunk
constexpr int kStepWaves = 4;  // waves per workgroup of the fused pass
enum StepMode { MODE_NONE = 0, MODE_DBL = 1, MODE_LR = 2, MODE_DBL_SEG = 3, MODE_SRC = 4,
                MODE_DUAL = 5, MODE_DBLA = 6, MODE_DUALP = 7 };
// flag bits of the fused pass: rows longer than the gather slots exist; the
// doubling modes may take the own-row words from the diagonal's gather slot
// (kpm_csr::diag_slots <= slots)
constexpr int kStepTail = 1;
constexpr int kStepDiagGather = 2;
// ... and (MODE_NONE) write the split-BF16 contraction planes (bf16_planes.h)
// of out to out2, with kStepPlanesOnly instead of out; MODE_DUALP writes the
// dual pass's second product as planes
constexpr int kStepPlanes = 4;
constexpr int kStepPlanesOnly = 8;

// One fused pass out <- alpha*A*x + beta*z with optional dot partials:
//   MODE_DBL: d0 = <x|x>, d1 = <x|out>;  MODE_LR: d0 = <l|out>.
// z may alias out (in place) and may be null when beta == 0.
struct PendingCombine;
struct StepLaunch {
  const kpm_csr* A;
  const kpm_block* x;
  const kpm_block* z;
  kpm_block* out;
  const kpm_block* l;
  double alpha, beta;
  int mode;
  double gamma = 0.0;                     // MODE_SRC, MODE_DUAL only
  double2* part_out = nullptr;            // dot partials (null: ctx->partials)
  int max_grid = 0;                       // with part_out: workgroups it holds (0: kMaxGridPerCU * CUs)
  const PendingCombine* pend = nullptr;   // previous step's combine, run in this launch
  const void* vals2 = nullptr;            // MODE_DUAL: second operator's values (complex)
  kpm_block* out2 = nullptr;              // MODE_DUAL: second output
  // MODE_DUAL: out2 split into blocks of out2_cols complex columns, each an
  // NH x out2_cols array, out2_stride complex values apart (0: out's layout)
  int64_t out2_cols = 0, out2_stride = 0;
  // split-BF16 contraction planes of this pass's contraction operand (out in
  // MODE_NONE, the second product in MODE_DUAL), one slot at this address
  void* planes = nullptr;
  bool planes_only = false;  // MODE_NONE: write the planes instead of out
};
// Word placement of the dual pass's second output (words of the block type):
// word w of a row goes to block w / cw, qs words apart, at row * rw + w % cw.
struct SplitOut {
  int rw = 0, cw = 1;
  int64_t qs = 0;
};
// Launch geometry picked by the tuner for (LPR, rows per lane, grid size).
struct StepConfig {
  int unroll = 2;     // rows per lane per tile
  int grid = 0;       // persistent workgroups (multiple of 8); 0 = auto
  int chunk = 0;      // > 0: one-shot schedule, this many tiles per workgroup
  int xmap = 0;       // one-shot schedule: 1 = chunks grouped per XCD, m >= 2 = runs of m chunks per XCD
};
int launch_step(kpm_ctx* ctx, const StepLaunch& s, const StepConfig& cfg, int* grid_used);
// geometry of the fused pass for a block shape: lanes per row, row-chunks,
// rows per tile, nonzeros per row in flight, tail flag
struct StepGeom {
  int RW, lpr, chunks, unr, kb, has_tail;
  int64_t TR;
};
StepGeom step_geom(const kpm_csr* A, int64_t NR, int prec, const StepConfig& cfg);
// persistent grid for ntiles row tiles (multiple of 8, up to 8 workgroups per CU)
int default_grid(kpm_ctx* ctx, int64_t ntiles);
// segmented (batched) MODE_DBL pass on G workgroups with per-workgroup tile
// ranges tile_lo[G+1], tile realisations treal[ntiles], segment bases seg0[G]
int launch_step_seg(kpm_ctx* ctx, const StepLaunch& s, const StepGeom& g, int G, int64_t ntiles,
                    const int* tile_lo, const int* treal, const int* seg0, double2* partials);
// dot-only pass: d0 = <a|b> per column, partials into ctx->partials
int launch_dot(kpm_ctx* ctx, const kpm_block* a, const kpm_block* b, int* grid_used);
// Reduce partials [grid][nd][NRp] in a fixed order and write moment columns:
//   mu[:, tcol[d]] = alpha[d]*sum_d + bscale[d]*mu[:, bcol[d]]  (bcol < 0: none)
// (struct Combine is declared at the top of this header)
int launch_combine(kpm_ctx* ctx, int grid, int64_t NR, int64_t NRp, const Combine& c,
                   void* mu_dev, int64_t NCmu);
// the same reduction over an explicit partials buffer
int launch_combine_from(kpm_ctx* ctx, const double2* partials, int grid, int64_t NR, int64_t NRp,
                        const Combine& c, void* mu_dev);
// A combine deferred into the next recurrence launch of the same engine: the
// first nd*NR workgroups of that launch reduce these partials (written by the
// previous launch, complete at the kernel boundary) before their own tiles,
// which saves one launch per step.  Dot steps then alternate between two
// partials buffers of at most kMaxGridPerCU * CUs workgroups each.
// Large one-shot grids reduce in two levels: the launch after a step sums
// its partials in slices of kSlice workgroups (first level, l1_*), and the
// launch after that reduces the slice sums into the moment columns (part,
// G = slices).  Both levels add in a fixed order.
constexpr int kSlice = 1024;
constexpr int kTwoLevelMin = 4096;  // grids above this reduce in two levels
struct PendingCombine {
  const double2* part = nullptr;
  double2* mu = nullptr;
  Combine cb{0, {0, 0}, {0.0, 0.0}, {-1, -1}, {0.0, 0.0}};
  int G = 0, NRp = 0, NR = 0;
  const double2* l1_part = nullptr;  // first level: partials of l1_G workgroups
  double2* l1_out = nullptr;         // -> slice sums [slice][d][col]
  int l1_G = 0, l1_nd = 0;
};
// first-level reduction as a launch of its own (the engines' last steps)
int launch_reduce1(kpm_ctx* ctx, const double2* partials, int G, int nd, int64_t NR, int64_t NRp,
                   double2* out);
constexpr int kMaxGridPerCU = 16;
// one-shot (chunked) grids may use up to this many workgroups per CU; the
// engines size their partials buffers for the tuned grid of the shape
constexpr int kMaxChunkGridPerCU = 128;
// workgroups the partials buffers of the engines must hold for a shape:
// kMaxGridPerCU * CUs, or the tuned one-shot grid when it is larger
int partials_grid(kpm_ctx* ctx, const kpm_csr* A, int64_t NH, int64_t NR, int prec, int mode);
int partials_grid_cfg(kpm_ctx* ctx, const kpm_csr* A, int64_t NH, int64_t NR, int prec,
                      const StepConfig& cfg);
// ms of `reps` doubling steps of the kpm_1d engine (fused passes and their
// deferred reductions) under launch shape cfg on blocks X (v_n) and Z
// (v_{n-1}, overwritten), after two untimed steps: what the tuner compares
int time_engine_steps(kpm_ctx* ctx, const kpm_csr* H, kpm_block* X, kpm_block* Z,
                      const StepConfig& cfg, int reps, float* ms, int mode = MODE_DBL);
// layout conversions between the host column-major ComplexF64 block (in the
// staging buffer, already on the device) and a device recurrence block
int launch_upload(kpm_ctx* ctx, const void* staging, kpm_block* b);
int launch_download(kpm_ctx* ctx, const kpm_block* b, void* staging);

// streaming copy of nwords 16-byte words (calibration / bandwidth ceiling)
int launch_copy16(kpm_ctx* ctx, const void* src, void* dst, int64_t nwords);

// ---- contraction (gemm.hip) -----------------------------------------------------
// mu_blk[n][m] (n < nR, m < nL) = scale * sum_k conj(L_m[k]) * R_n[k], over
// K = NH*NRp complex entries per slot, accumulated into the host-layout
// table mu (NC x NC column-major device double2) at (n0 + n, m0 + m) with the
// parity mask (0 none, 1 keep odd n+m, 2 keep even; 1-based n, m as in the
// reference).
// K: complex values contracted per slot; pitch: complex values between slots
int launch_contract(kpm_ctx* ctx, const void* Lslots, int nL, const void* Rslots, int nR,
                    int64_t K, int64_t pitch, int prec, int gemm_mode, double scale, void* mu_dev,
                    int64_t NC, int64_t n0, int64_t m0, int parity);
// the split-BF16 contraction of slots stored as planes (bf16_planes.h; 2 *
// pitch words per slot), for 128 or more slots a side: bitwise the table the
// split-BF16 mode of launch_contract gives from the slots they were split from
int launch_contract_planes(kpm_ctx* ctx, const void* Lplanes, int nL, const void* Rplanes, int nR,
                           int64_t K, int64_t pitch, double scale, void* mu_dev, int64_t NC, int64_t n0,
                           int64_t m0, int parity);

// ---- reverse mode (adjoint.hip) ---------------------------------------------------
// window of reverse steps whose gradient terms one SDDMM pass adds
constexpr int kWin = 8;
// M[k] += sum_{w < W} c[w] sum_r conj(beta_w[a, r]) alpha_w[colind[k], r]
// for every stored entry k (row a) of A; W <= kWin.  Rows wider than 64
// words run in sddmm_chunks() row chunks, chunk c adding into
// M + c * mstride (mstride >= nnz; the caller sums the chunks in order).
int sddmm_chunks(int64_t NR, int prec);
int launch_sddmm(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* const* beta,
                 const kpm_block* const* alpha, const double* c, int W, double2* M,
                 size_t mstride);
// out = alpha*A*x + beta*z (+ gamma*src in MODE_SRC), no dots
int adj_step(kpm_ctx* ctx, const kpm_csr* A, const kpm_block* x, const kpm_block* z,
             kpm_block* out, const kpm_block* src, double alpha, double beta, double gamma, int mode);

// looks up the scratch array's device address for ctx->device (noalloc.hip)
int noalloc_scratch_init(kpm_ctx* ctx);

// ---- kpm_1d! doubling engine (engine.hip) --------------------------------------------
// w0 holds psi on entry (overwritten), w1 is scratch; mu_dev (NR x NC
// double2, every column written) receives the moments.  part0 / part1: two
// dot-partial buffers of max_grid * 2 * NRp double2 each (null: the
// context's own, grown as needed).  Launches only, no allocation or
// synchronisation when both partial buffers are given.
int run_1d_doubling_into(kpm_ctx* ctx, const kpm_csr* H, int64_t NC, kpm_block* w0, kpm_block* w1,
                         void* mu_dev, double2* part0, double2* part1, int max_grid);

// tuner table (tune.hip): best StepConfig for a shape, or defaults; record one
StepConfig tuned_config(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, int mode);
void tune_record(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, const StepConfig& c);
// the same for the probe-summed doubling pass (MODE_DBLA) alone
void tune_record_avg(kpm_ctx* ctx, const kpm_csr* A, int64_t NR, int prec, const StepConfig& c);

}  // namespace kpm
