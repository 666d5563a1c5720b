// Batched AMIS over many independent posteriors in one persistent launch:
// data structures shared by the host side (batch.hip) and the kernel
// translation units (batch_k.hip).
//
// A batch is a vector of prepared AMIS samplers, each with the meaning of
// importance_sample!(prepare_sampler(rng, target, AMIS(q0; rounds, round_size)))
// (src/methods/amis.jl:42-47, src/adaptive_gaussian_execution.jl:897-1112) for a
// built-in regression target (src/targets.jl:17-43 batch contract).  One
// workgroup owns one job for all of its rounds; workgroups pull jobs from a
// device work queue in the order the host packer chose.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace issampler {

constexpr int kBatchMaxDim = 32;
constexpr int kBatchStage = kBatchMaxDim + 1;  // row stride (elements) of a wave's 64-sample LDS stage

// Batch slot layout (FP64 master, T copy and LDS copy share it), for the
// job's padded dimension P = d rounded up to 4:
//   [ mu (P) | 1/L_rr (P) | L rows 0..P-1 | lognorm | logmass ],
// row r of L holds r + 1 entries padded to a multiple of 4, so every row starts
// 16-byte aligned and the density kernels read it with wide LDS loads.  Rows
// d..P-1 are padding: mu 0, L row 0, 1/L_rr 1.
__host__ __device__ constexpr int bslot_rowoff(int r) {
  return 8 * (r / 4) * (r / 4 + 1) + 4 * (r % 4) * (r / 4 + 1);
}
__host__ __device__ constexpr int bslot_pad(int d) { return (d + 3) & ~3; }
__host__ __device__ constexpr int bslot_lognorm(int P) { return 2 * P + bslot_rowoff(P); }
__host__ __device__ constexpr int bslot_size(int P) { return (bslot_lognorm(P) + 2 + 3) & ~3; }

// FP32 jobs keep, per LDS slot, the inverse factor M = L^{-1} (FP64 solve, FP32 values)
// as v_mfma_f32_16x16x4f32 A operands for the whitened coordinates y = M (x - mu):
// row tiles i < ceil(P / 16), K steps kk < min(4 (i + 1), P / 4) (M is lower
// triangular and coordinates past P are not read), 64 lanes per (i, kk) pair, lane l
// holding M[16 i + l % 16][4 kk + l / 16] (rows past P zero).
__host__ __device__ constexpr int bfrag_pairs(int P) {
  int n = 0;
  for (int i = 0; i < (P + 15) / 16; ++i) n += (4 * (i + 1) < P / 4) ? 4 * (i + 1) : P / 4;
  return n;
}
__host__ __device__ constexpr int bfrag_elems(int P) { return 64 * bfrag_pairs(P); }

// Samples are stored in blocks of 64: coordinate c of sample i sits at
// bx_base(i, P) + 64 c, so a wave's 64 samples of one coordinate are one
// contiguous 256-byte (FP32) segment and every coordinate is a constant offset.
__host__ __device__ inline int64_t bx_base(int64_t i, int P) { return (i >> 6) * (64 * (int64_t)P) + (i & 63); }

// Per-job descriptor in device memory (read by the kernel with scalar loads).
template <class T>
struct BatchJob {
  int d, dp, rounds, student, tkind, slot_elems, rs, us;  // dp: padded dimension P
  int tbf16;            // FP32 linear regression, P >= 8: the centred BF16 target
  int64_t N, cap, row_tiles, n_rows;
  double dof, logconstant, inv_two_var, inv_two_prior, tconst;
  T logN, fdof, fhalf, pad_;
  uint64_t key0, key1;
  T* X;                 // samples in 64-sample blocks: X[(i / 64) 64 d + 64 c + i % 64]
  T* lt;                // log targets
  T* ln;                // AMIS log numerators
  T* lw;                // log weights
  const T* xpack;       // data rows in MFMA A-fragment order (row_tiles x P/4 x 64)
  const T* ypack;       // responses in MFMA output order (row_tiles x 64 x 4); with tbf16 the
                        // centred residuals y - X x0 (FP64, rounded to T), one per row
  const uint32_t* xbf;  // tbf16: rows split x = xh + xl in BF16 as 16x16x32 A operands,
                        // [row tile][hi, lo][lane][8 BF16]
  const double* x0;     // tbf16: the centre (P entries, zero past d)
  double* slots64;      // rounds + 1 slots, FP64
  T* slotsT;            // the same slots in T
  double* committed64;  // committed proposal (slot 0 of the next run)
  T* committedT;
  double* round_ess;
  double* round_logz;
  int* fit_status;              // per round: 0 ok, 1 all-zero weights, 2 not positive definite, 3 invalid
  unsigned long long* fail;     // 3 words per round (src/execution.jl:669-699)
};

template <class T>
struct BatchArgs {
  const BatchJob<T>* jobs;
  const int32_t* order;         // work-queue position -> job index
  int njobs;
  unsigned int* queue;          // work-queue head (zeroed by k_batch_begin)
  const unsigned long long* run_hi;  // device word: run index << 32
  int slots_off;                // element offset of the slot area in LDS (after the wave stages)
};

// LDS bytes of one launch: W wave stages plus the largest job's rounds x slot
// (for FP32 jobs, with the slots' factor fragments after them).
inline size_t batch_lds_bytes(int W, int tsize, int max_slot_area) {
  return (size_t)W * 64 * kBatchStage * tsize + (size_t)max_slot_area * tsize;
}

// Kernel launchers, one translation unit per (T, W) (batch_k.hip).
template <class T, int W>
void launch_batch_amis(const BatchArgs<T>& a, int blocks, size_t lds, hipStream_t s);
template <class T>
void launch_batch_begin(const BatchJob<T>* jobs, int njobs, unsigned int* queue, hipStream_t s);
template <class T>
void launch_batch_end(const BatchJob<T>* jobs, int njobs, int commit, int* status, unsigned long long* run_word,
                      hipStream_t s);

}  // namespace issampler
