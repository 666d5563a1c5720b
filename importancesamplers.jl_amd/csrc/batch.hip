// Host side of the batched AMIS engine: job validation and packing, device
// pools, the work-queue order, graph-captured runs, and the is_batch_* C ABI
// declared in include/issampler.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "../../include/issampler.h"
#include "batch.h"
#include "host.h"
#include "kernels.h"

using namespace issampler;
using namespace issampler::host;

namespace {

constexpr int kBatchMaxRounds = 64;

struct BatchJobHost {
  int d = 0, P = 0, rounds = 0, family = 0, tkind = 0, SS = 0;
  int64_t N = 0, cap = 0, n_rows = 0, row_tiles = 0;
  double dof = 0.0, logconstant = 0.0, inv_two_var = 0.0, inv_two_prior = 0.0, tconst = 0.0;
  RngKey key{0, 0};
  // element offsets into the pools
  size_t x_off = 0, log_off = 0, xpack_off = 0, ypack_off = 0, slot_off = 0, com_off = 0, round_off = 0;
  size_t xbf_off = 0, x0_off = 0;
  bool bf16 = false;  // the centred BF16 target (FP32 linear regression, P >= 8)
  double cost = 0.0;
  std::vector<double> slot0;  // initial proposal, batch layout (FP64)
  std::vector<double> x0;     // the BF16 target's centre: the initial proposal's location
};

// Packs a proposal into the batch slot layout for padded dimension P
// (batch.h): mu, 1/L_rr, rows of L with r + 1 entries padded to 4, lognorm.
void pack_bslot(const HostProp& p, int d, int P, double* slot) {
  std::fill(slot, slot + bslot_size(P), 0.0);
  for (int r = 0; r < P; ++r) {
    slot[P + r] = 1.0;
    if (r >= d) continue;
    slot[r] = p.mu[r];
    slot[P + r] = slot_inv_entry<double>(p.L[r + (size_t)r * d]);
    for (int c = 0; c <= r; ++c) slot[2 * P + bslot_rowoff(r) + c] = p.L[r + (size_t)c * d];
  }
  slot[bslot_lognorm(P)] = p.lognorm;
}

void unpack_bslot(const double* slot, int d, int P, double* mu, double* L, double* lognorm) {
  for (int r = 0; r < d; ++r) {
    if (mu) mu[r] = slot[r];
    if (L)
      for (int c = 0; c < d; ++c) L[r + (size_t)c * d] = c <= r ? slot[2 * P + bslot_rowoff(r) + c] : 0.0;
  }
  if (lognorm) *lognorm = slot[bslot_lognorm(P)];
}

}  // namespace

struct is_batch {
  is_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  std::atomic<bool> busy{false};
  int precision = IS_FP32;
  int njobs = 0;
  std::vector<BatchJobHost> jobs;
  std::vector<int32_t> order;
  int waves = 4, blocks_per_cu = 1, cus = 256;
  int max_slot_area = 0;
  size_t tsize = 4;
  DevBuf jobs_dev, order_dev, queue, run_dev, status_dev;
  DevBuf X, lt, ln, lw, xpack, ypack, slots64, slotsT, com64, comT, ress, rlogz, fit, fail;
  DevBuf xbf, x0d;  // the centred BF16 target's split rows and centres
  uint64_t run_index = 0;
  bool graph_dirty = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int timing = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  bool have_results = false;
  std::vector<int> status;
};

namespace {

struct BatchBusy {
  is_batch* b;
  bool ok;
  explicit BatchBusy(is_batch* bb) : b(bb) {
    bool expected = false;
    ok = b->busy.compare_exchange_strong(expected, true);
  }
  ~BatchBusy() {
    if (ok) b->busy.store(false);
  }
};

template <class T>
void pack_data(const BatchJobHost& J, const double* X, const double* y, T* xpack, T* ypack) {
  // A fragment of row tile rt, K step k, lane l: X[16 rt + l % 16][4 k + l / 16]
  const int KS = J.P / 4;
  for (int64_t rt = 0; rt < J.row_tiles; ++rt)
    for (int k = 0; k < KS; ++k)
      for (int l = 0; l < 64; ++l) {
        const int64_t r = 16 * rt + (l & 15);
        const int c = 4 * k + (l >> 4);
        xpack[(rt * KS + k) * 64 + l] = (r < J.n_rows && c < J.d) ? (T)X[r * J.d + c] : T(0);
      }
  // responses in the MFMA output order: lane l, register q of row tile rt
  for (int64_t rt = 0; rt < J.row_tiles; ++rt)
    for (int l = 0; l < 64; ++l)
      for (int q = 0; q < 4; ++q) {
        const int row_in = sizeof(T) == 4 ? 4 * (l >> 4) + q : 4 * q + (l >> 4);
        const int64_t r = 16 * rt + row_in;
        ypack[(rt * 64 + l) * 4 + q] = r < J.n_rows ? (T)y[r] : T(0);
      }
}

// The centred BF16 linear-regression target's data: each row split as x = xh + xl
// (BF16 each, xl the rounding of x - xh) in the 16x16x32 A-operand order (lane l:
// row 16 rt + l % 16, coordinates 8 (l / 16) .. + 7), and the residuals at the centre,
// y - X x0 in FP64 rounded to FP32, one per row (16 per row tile; lane l reads the
// four of its output rows 4 (l / 16) .. + 3).
void pack_bf16(const BatchJobHost& J, const double* X, const double* y, uint16_t* xbf, float* ypack) {
  for (int64_t rt = 0; rt < J.row_tiles; ++rt)
    for (int l = 0; l < 64; ++l) {
      const int64_t r = 16 * rt + (l & 15);
      for (int e = 0; e < 8; ++e) {
        const int c = 8 * (l >> 4) + e;
        uint16_t h = 0, lo = 0;
        if (r < J.n_rows && c < J.d) {
          const double x = X[r * J.d + c];
          h = bf16_bits((float)x);
          lo = bf16_bits((float)(x - bf16_value(h)));
        }
        xbf[((rt * 2 + 0) * 64 + l) * 8 + e] = h;
        xbf[((rt * 2 + 1) * 64 + l) * 8 + e] = lo;
      }
    }
  for (int64_t r = 0; r < 16 * J.row_tiles; ++r) {
    double v = 0.0;
    if (r < J.n_rows) {
      v = y[r];
      for (int c = 0; c < J.d; ++c) v -= X[r * J.d + c] * J.x0[c];
    }
    ypack[r] = (float)v;
  }
}

template <class T>
int batch_upload(is_batch* b, const std::vector<const double*>& Xs, const std::vector<const double*>& ys) {
  size_t nx = 0, nlog = 0, nxp = 0, nyp = 0, nslot = 0, ncom = 0, nround = 0, nbf = 0, nx0 = 0;
  for (auto& J : b->jobs) {
    J.x_off = nx;
    nx += (size_t)(J.cap / 64) * 64 * J.P;
    J.log_off = nlog;
    nlog += (size_t)J.cap;
    J.xpack_off = nxp;
    if (!J.bf16) nxp += (size_t)J.row_tiles * (J.P / 4) * 64;
    J.xbf_off = nbf;
    if (J.bf16) nbf += (size_t)J.row_tiles * 2 * 64 * 8;
    J.x0_off = nx0;
    if (J.bf16) nx0 += (size_t)J.P;
    J.ypack_off = nyp;
    nyp += (size_t)J.row_tiles * (J.bf16 ? 16 : 64 * 4);
    J.slot_off = nslot;
    nslot += (size_t)(J.rounds + 1) * J.SS;
    J.com_off = ncom;
    ncom += (size_t)J.SS;
    J.round_off = nround;
    nround += (size_t)J.rounds;
  }
  int rc;
  if ((rc = b->X.alloc(nx * sizeof(T)))) return rc;
  if ((rc = b->lt.alloc(nlog * sizeof(T)))) return rc;
  if ((rc = b->ln.alloc(nlog * sizeof(T)))) return rc;
  if ((rc = b->lw.alloc(nlog * sizeof(T)))) return rc;
  if ((rc = b->xpack.alloc(nxp * sizeof(T)))) return rc;
  if ((rc = b->ypack.alloc(nyp * sizeof(T)))) return rc;
  if ((rc = b->xbf.alloc(std::max<size_t>(nbf, 1) * sizeof(uint16_t)))) return rc;
  if ((rc = b->x0d.alloc(std::max<size_t>(nx0, 1) * sizeof(double)))) return rc;
  if ((rc = b->slots64.alloc(nslot * sizeof(double)))) return rc;
  if ((rc = b->slotsT.alloc(nslot * sizeof(T)))) return rc;
  if ((rc = b->com64.alloc(ncom * sizeof(double)))) return rc;
  if ((rc = b->comT.alloc(ncom * sizeof(T)))) return rc;
  if ((rc = b->ress.alloc(nround * sizeof(double)))) return rc;
  if ((rc = b->rlogz.alloc(nround * sizeof(double)))) return rc;
  if ((rc = b->fit.alloc(nround * sizeof(int)))) return rc;
  if ((rc = b->fail.alloc(3 * nround * sizeof(uint64_t)))) return rc;
  if ((rc = b->status_dev.alloc((size_t)b->njobs * sizeof(int)))) return rc;
  if ((rc = b->queue.alloc(64))) return rc;
  if ((rc = b->run_dev.alloc(64))) return rc;
  HIP_TRY(locked_memset(b->run_dev.p, 0, 64));
  HIP_TRY(locked_memset(b->slots64.p, 0, b->slots64.bytes));
  HIP_TRY(locked_memset(b->slotsT.p, 0, b->slotsT.bytes));
  // data in MFMA fragment order, committed slots, job table
  std::vector<T> xp(nxp), yp(nyp);
  std::vector<uint16_t> xb(nbf);
  std::vector<double> x0s(nx0, 0.0);
  std::vector<double> c64(ncom);
  std::vector<BatchJob<T>> table(b->njobs);
  for (int j = 0; j < b->njobs; ++j) {
    const BatchJobHost& H = b->jobs[j];
    if constexpr (sizeof(T) == 4) {
      if (H.bf16) {
        pack_bf16(H, Xs[j], ys[j], xb.data() + H.xbf_off, yp.data() + H.ypack_off);
        std::copy(H.x0.begin(), H.x0.end(), x0s.begin() + H.x0_off);
      }
    }
    if (!H.bf16) {
      pack_data<T>(H, Xs[j], ys[j], xp.data() + H.xpack_off, yp.data() + H.ypack_off);
    }
    std::copy(H.slot0.begin(), H.slot0.end(), c64.begin() + H.com_off);
    BatchJob<T>& J = table[j];
    std::memset(&J, 0, sizeof(J));
    J.d = H.d;
    J.dp = H.P;
    J.rounds = H.rounds;
    J.student = H.family == IS_STUDENT_T;
    J.tkind = H.tkind == IS_TARGET_LOGISTIC ? kTLogistic : kTLinreg;
    J.tbf16 = H.bf16 ? 1 : 0;
    J.slot_elems = H.SS;
    J.rs = radial_normal_stride(H.family, H.dof);
    J.us = radial_uniform_stride(H.family, H.dof);
    J.N = H.N;
    J.cap = H.cap;
    J.row_tiles = H.row_tiles;
    J.n_rows = H.n_rows;
    J.dof = H.dof;
    J.logconstant = H.logconstant;
    J.inv_two_var = H.inv_two_var;
    J.inv_two_prior = H.inv_two_prior;
    J.tconst = H.tconst;
    J.logN = (T)std::log((double)H.N);
    J.fdof = (T)H.dof;
    J.fhalf = (T)((H.dof + (double)H.d) / 2.0);
    J.key0 = H.key.k0;
    J.key1 = H.key.k1;
    J.X = b->X.as<T>() + H.x_off;
    J.lt = b->lt.as<T>() + H.log_off;
    J.ln = b->ln.as<T>() + H.log_off;
    J.lw = b->lw.as<T>() + H.log_off;
    J.xpack = b->xpack.as<T>() + H.xpack_off;
    J.ypack = b->ypack.as<T>() + H.ypack_off;
    J.xbf = H.bf16 ? reinterpret_cast<const uint32_t*>(b->xbf.as<uint16_t>() + H.xbf_off) : nullptr;
    J.x0 = H.bf16 ? b->x0d.as<double>() + H.x0_off : nullptr;
    J.slots64 = b->slots64.as<double>() + H.slot_off;
    J.slotsT = b->slotsT.as<T>() + H.slot_off;
    J.committed64 = b->com64.as<double>() + H.com_off;
    J.committedT = b->comT.as<T>() + H.com_off;
    J.round_ess = b->ress.as<double>() + H.round_off;
    J.round_logz = b->rlogz.as<double>() + H.round_off;
    J.fit_status = b->fit.as<int>() + H.round_off;
    J.fail = b->fail.as<unsigned long long>() + 3 * H.round_off;
  }
  std::vector<T> cT(c64.begin(), c64.end());
  for (int j = 0; j < b->njobs; ++j) {  // inverse diagonals that are not normal in T: NaN
    const BatchJobHost& H = b->jobs[j];
    for (int r = 0; r < H.d; ++r) cT[H.com_off + H.P + r] = slot_inv_value<T>(c64[H.com_off + H.P + r]);
  }
  HIP_TRY(locked_memcpy(b->xpack.p, xp.data(), nxp * sizeof(T), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(b->ypack.p, yp.data(), nyp * sizeof(T), hipMemcpyHostToDevice));
  if (nbf) HIP_TRY(locked_memcpy(b->xbf.p, xb.data(), nbf * sizeof(uint16_t), hipMemcpyHostToDevice));
  if (nx0) HIP_TRY(locked_memcpy(b->x0d.p, x0s.data(), nx0 * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(b->com64.p, c64.data(), ncom * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(locked_memcpy(b->comT.p, cT.data(), ncom * sizeof(T), hipMemcpyHostToDevice));
  if ((rc = b->jobs_dev.alloc(table.size() * sizeof(BatchJob<T>)))) return rc;
  HIP_TRY(locked_memcpy(b->jobs_dev.p, table.data(), table.size() * sizeof(BatchJob<T>), hipMemcpyHostToDevice));
  return IS_OK;
}

// Work-queue order: longest estimated job first (LPT list scheduling), so the
// persistent workgroups finish together.  Cost per round sample: the MFMA
// target over the data rows, the mixture and append densities, the draw.
void lpt_order(is_batch* b) {
  for (auto& J : b->jobs) {
    const double P = J.P, mix = 0.5 * P * P * (J.rounds + 1);
    J.cost = (double)J.rounds * J.N * (32.0 * J.row_tiles * (P / 4.0) + mix + 64.0 * P);
  }
  b->order.resize(b->njobs);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int a, int c) { return b->jobs[a].cost > b->jobs[c].cost; });
}

int upload_order(is_batch* b) {
  int rc = b->order_dev.alloc((size_t)b->njobs * sizeof(int32_t));
  if (rc) return rc;
  HIP_TRY(locked_memcpy(b->order_dev.p, b->order.data(), (size_t)b->njobs * sizeof(int32_t), hipMemcpyHostToDevice));
  b->graph_dirty = true;
  return IS_OK;
}

size_t batch_lds(const is_batch* b, int W) { return batch_lds_bytes(W, (int)b->tsize, b->max_slot_area); }

template <class T>
void launch_amis(is_batch* b, int W, int blocks, hipStream_t s) {
  BatchArgs<T> a{};
  a.jobs = b->jobs_dev.as<BatchJob<T>>();
  a.order = b->order_dev.as<int32_t>();
  a.njobs = b->njobs;
  a.queue = b->queue.as<unsigned int>();
  a.run_hi = b->run_dev.as<unsigned long long>();
  a.slots_off = W * 64 * kBatchStage;
  const size_t lds = batch_lds(b, W);
  if (W == 2)
    launch_batch_amis<T, 2>(a, blocks, lds, s);
  else if (W == 8)
    launch_batch_amis<T, 8>(a, blocks, lds, s);
  else
    launch_batch_amis<T, 4>(a, blocks, lds, s);
}

// begin, persistent kernel, end: one run's launch sequence
template <class T>
void batch_sequence(is_batch* b, int W, int blocks, int commit, bool timed, hipStream_t s) {
  const BatchJob<T>* jobs = b->jobs_dev.as<BatchJob<T>>();
  launch_batch_begin<T>(jobs, b->njobs, b->queue.as<unsigned int>(), s);
  if (timed) (void)hipEventRecord(b->ev0, s);
  launch_amis<T>(b, W, blocks, s);
  if (timed) (void)hipEventRecord(b->ev1, s);
  launch_batch_end<T>(jobs, b->njobs, commit, b->status_dev.as<int>(), b->run_dev.as<unsigned long long>(), s);
}

template <class T>
int batch_run_typed(is_batch* b, int W, int bpc, int commit, bool eager) {
  hipStream_t s = b->stream;
  const int blocks = std::max(1, std::min(b->njobs, bpc * b->cus));
  const bool timed = b->timing || eager;
  if (!timed) {
    if (b->graph_dirty || !b->graph_exec) {
      CaptureGuard cg(capture_lock());
      if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
      if (b->graph) (void)hipGraphDestroy(b->graph);
      b->graph_exec = nullptr;
      b->graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      batch_sequence<T>(b, W, blocks, commit, false, s);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (e != hipSuccess) return fail(IS_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
      b->graph = g;
      HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
      b->graph_dirty = false;
    }
    HIP_TRY(hipGraphLaunch(b->graph_exec, s));
  } else {
    batch_sequence<T>(b, W, blocks, commit, true, s);
    HIP_TRY(hipGetLastError());
  }
  b->status.assign(b->njobs, 0);
  HIP_TRY(hipMemcpyAsync(b->status.data(), b->status_dev.p, (size_t)b->njobs * sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (timed) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess) b->last_ms = ms;
  }
  if (commit) b->run_index += 1;
  b->have_results = true;
  for (int st : b->status)
    if (st) return fail(IS_EFAILED, "one or more jobs of the batch failed (see is_batch_job_failure)");
  return IS_OK;
}

int batch_run(is_batch* b, int W, int bpc, int commit, bool eager) {
  HIP_TRY(hipSetDevice(b->ctx->device));
  return b->precision == IS_FP32 ? batch_run_typed<float>(b, W, bpc, commit, eager)
                                 : batch_run_typed<double>(b, W, bpc, commit, eager);
}

bool tuning_fits(const is_batch* b, int W, int bpc) {
  if (W != 2 && W != 4 && W != 8) return false;
  if (bpc < 1 || bpc * W > 32) return false;
  return batch_lds(b, W) * (size_t)bpc <= (size_t)160 * 1024;
}

template <class T>
int copy_job_T(is_batch* b, const DevBuf& pool, size_t off, int64_t n, double* out) {
  std::vector<T> h((size_t)n);
  HIP_TRY(locked_memcpy(h.data(), pool.as<T>() + off, (size_t)n * sizeof(T), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; ++i) out[i] = (double)h[(size_t)i];
  return IS_OK;
}

#define BATCH_JOB_CHECK(b, job)                                                         \
  do {                                                                                  \
    if (!(b) || (job) < 0 || (job) >= (b)->njobs) return fail(IS_EINVAL, "bad job index"); \
  } while (0)

}  // namespace

// ===================================================================== C ABI
extern "C" {

int is_batch_create(is_ctx* ctx, const is_job_desc* descs, int32_t njobs, int32_t precision, is_batch** out) {
  if (!ctx || !descs || !out || njobs < 1) return fail(IS_EINVAL, "null argument or empty batch");
  if (precision != IS_FP32 && precision != IS_FP64) return fail(IS_EINVAL, "unknown precision");
  HIP_TRY(hipSetDevice(ctx->device));
  auto b = new is_batch();
  std::unique_ptr<is_batch> guard(b);
  b->ctx = ctx;
  b->precision = precision;
  b->tsize = precision == IS_FP32 ? 4 : 8;
  b->njobs = njobs;
  b->jobs.resize(njobs);
  std::vector<const double*> Xs(njobs), ys(njobs);
  for (int j = 0; j < njobs; ++j) {
    const is_job_desc& D = descs[j];
    BatchJobHost& H = b->jobs[j];
    if (D.method != IS_AMIS) return fail(IS_EUNSUPPORTED, "job %d: the batch runs AMIS jobs", j);
    if (D.transform_kinds) return fail(IS_EUNSUPPORTED, "job %d: transforms are not supported in a batch", j);
    if (D.target_kind != IS_TARGET_LINREG && D.target_kind != IS_TARGET_LOGISTIC)
      return fail(IS_EUNSUPPORTED, "job %d: the batch runs regression targets", j);
    if (D.dim < 1 || D.dim > kBatchMaxDim) return fail(IS_EUNSUPPORTED, "job %d: dimension outside [1, %d]", j, kBatchMaxDim);
    if (D.rounds < 1 || D.rounds > kBatchMaxRounds || !D.round_size)
      return fail(IS_EINVAL, "job %d: rounds must be in [1, %d]", j, kBatchMaxRounds);
    for (int t = 1; t < D.rounds; ++t)
      if (D.round_size[t] != D.round_size[0]) return fail(IS_EUNSUPPORTED, "job %d: the batch needs equal round sizes", j);
    if (D.round_size[0] < 1) return fail(IS_EINVAL, "job %d: round sizes must be positive", j);
    if (D.family != IS_GAUSSIAN && D.family != IS_STUDENT_T) return fail(IS_EINVAL, "job %d: unknown family", j);
    if (D.family == IS_STUDENT_T && !(std::isfinite(D.dof) && D.dof > 2.0))
      return fail(IS_EINVAL, "job %d: covariance adaptation requires Student-t degrees of freedom > 2", j);
    if (!D.location || !D.scale) return fail(IS_EINVAL, "job %d: proposal location and scale are required", j);
    H.d = D.dim;
    H.P = bslot_pad(D.dim);
    H.rounds = D.rounds;
    H.family = D.family;
    H.dof = D.family == IS_STUDENT_T ? D.dof : 0.0;
    H.tkind = D.target_kind;
    H.N = D.round_size[0];
    H.cap = ((int64_t)H.rounds * H.N + 63) / 64 * 64;
    if ((int64_t)H.rounds * H.N >= ((int64_t)1 << 31)) return fail(IS_EUNSUPPORTED, "job %d: sample budget exceeds 2^31 - 1", j);
    H.SS = bslot_size(H.P);
    H.key = RngKey{D.seed, D.stream};
    if (H.family == IS_STUDENT_T) H.logconstant = student_t_logconstant(H.dof, H.d);
    HostProp p;
    std::string err = make_prop(H.family, H.dof, D.scale_kind, H.d, D.location, D.scale, p);
    if (!err.empty()) return fail(IS_EINVAL, "job %d: %s", j, err.c_str());
    H.slot0.assign(H.SS, 0.0);
    pack_bslot(p, H.d, H.P, H.slot0.data());
    H.bf16 = precision == IS_FP32 && D.target_kind == IS_TARGET_LINREG && H.P >= 8;
    if (H.bf16) {
      H.x0.assign(H.P, 0.0);
      for (int c = 0; c < H.d; ++c) H.x0[c] = p.mu[c];
    }
    // target: [n, p, prior_var, X, y] (logistic) or [n, p, sigma, tau, X, y] (linear regression)
    const int head = H.tkind == IS_TARGET_LOGISTIC ? 3 : 4;
    const double* tp = D.target_params;
    if (!tp || D.target_params_len < head) return fail(IS_EINVAL, "job %d: regression target parameters are missing", j);
    const double nd = tp[0], pd = tp[1];
    if (!(nd >= 1.0) || nd != std::floor(nd) || pd != (double)H.d)
      return fail(IS_EINVAL, "job %d: regression target needs n >= 1 rows and p == dim coefficients", j);
    H.n_rows = (int64_t)nd;
    if (D.target_params_len < head + H.n_rows * H.d + H.n_rows)
      return fail(IS_EINVAL, "job %d: regression data is shorter than n x p + n", j);
    H.row_tiles = (H.n_rows + 15) / 16;
    const double pi2 = 2.0 * 3.141592653589793;
    if (H.tkind == IS_TARGET_LOGISTIC) {
      const double pv = tp[2];
      if (!(pv > 0.0) || !std::isfinite(pv)) return fail(IS_EINVAL, "job %d: prior variance must be positive", j);
      H.inv_two_prior = 0.5 / pv;
      H.tconst = -0.5 * (double)H.d * std::log(pi2 * pv);
    } else {
      const double sigma = tp[2], tau = tp[3];
      if (!(sigma > 0.0) || !(tau > 0.0) || !std::isfinite(sigma) || !std::isfinite(tau))
        return fail(IS_EINVAL, "job %d: sigma and tau must be positive", j);
      H.inv_two_var = 0.5 / (sigma * sigma);
      H.inv_two_prior = 0.5 / (tau * tau);
      H.tconst = -(double)H.n_rows * std::log(sigma) - 0.5 * (double)H.n_rows * std::log(pi2) -
                 (double)H.d * std::log(tau) - 0.5 * (double)H.d * std::log(pi2);
    }
    Xs[j] = tp + head;
    ys[j] = tp + head + H.n_rows * H.d;
    // FP32: each LDS slot's factor fragments after the slots (element offset rounds x SS)
    const int frag_area = precision == IS_FP32 ? H.rounds * bfrag_elems(H.P) : 0;
    b->max_slot_area = std::max(b->max_slot_area, H.rounds * H.SS + frag_area);
  }
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
  b->cus = std::max(1, prop.multiProcessorCount);
  if (batch_lds(b, 2) > (size_t)160 * 1024) return fail(IS_EUNSUPPORTED, "the largest job's slots do not fit in LDS");
  b->waves = batch_lds(b, 4) <= (size_t)160 * 1024 ? 4 : 2;
  b->blocks_per_cu = 1;
  {
    CaptureGuard cg(capture_lock());
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  }
  HIP_TRY(hipEventCreate(&b->ev0));
  HIP_TRY(hipEventCreate(&b->ev1));
  int rc = precision == IS_FP32 ? batch_upload<float>(b, Xs, ys) : batch_upload<double>(b, Xs, ys);
  if (rc) return rc;
  lpt_order(b);
  if ((rc = upload_order(b))) return rc;
  *out = guard.release();
  return IS_OK;
}

int is_batch_destroy(is_batch* b) {
  if (!b) return IS_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  CaptureGuard cg(capture_lock());
  if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return IS_OK;
}

int is_batch_run(is_batch* b) {
  if (!b) return fail(IS_EINVAL, "null batch");
  BatchBusy g(b);
  if (!g.ok) return fail(IS_EBUSY, "batch is already running");
  return batch_run(b, b->waves, b->blocks_per_cu, 1, false);
}

int is_batch_set_tuning(is_batch* b, int32_t waves, int32_t blocks_per_cu) {
  if (!b) return fail(IS_EINVAL, "null batch");
  if (!tuning_fits(b, waves, blocks_per_cu))
    return fail(IS_EINVAL, "tuning (%d waves, %d blocks per CU) does not fit", waves, blocks_per_cu);
  if (waves != b->waves || blocks_per_cu != b->blocks_per_cu) b->graph_dirty = true;
  b->waves = waves;
  b->blocks_per_cu = blocks_per_cu;
  return IS_OK;
}

int is_batch_get_tuning(const is_batch* b, int32_t* waves, int32_t* blocks_per_cu) {
  if (!b) return fail(IS_EINVAL, "null batch");
  if (waves) *waves = b->waves;
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
  return IS_OK;
}

int is_batch_set_order(is_batch* b, const int32_t* order) {
  if (!b) return fail(IS_EINVAL, "null batch");
  if (!order) {
    lpt_order(b);
    return upload_order(b);
  }
  std::vector<char> seen(b->njobs, 0);
  for (int i = 0; i < b->njobs; ++i) {
    if (order[i] < 0 || order[i] >= b->njobs || seen[order[i]]) return fail(IS_EINVAL, "order is not a permutation");
    seen[order[i]] = 1;
  }
  b->order.assign(order, order + b->njobs);
  return upload_order(b);
}

int is_batch_get_order(const is_batch* b, int32_t* order) {
  if (!b || !order) return fail(IS_EINVAL, "null argument");
  std::copy(b->order.begin(), b->order.end(), order);
  return IS_OK;
}

int is_batch_trial(is_batch* b, int32_t waves, int32_t blocks_per_cu, double* kernel_ms) {
  if (!b) return fail(IS_EINVAL, "null batch");
  if (!tuning_fits(b, waves, blocks_per_cu))
    return fail(IS_EINVAL, "tuning (%d waves, %d blocks per CU) does not fit", waves, blocks_per_cu);
  BatchBusy g(b);
  if (!g.ok) return fail(IS_EBUSY, "batch is already running");
  const int rc = batch_run(b, waves, blocks_per_cu, 0, true);
  if (kernel_ms) *kernel_ms = b->last_ms;
  return rc;
}

int is_batch_set_timing(is_batch* b, int32_t enable) {
  if (!b) return fail(IS_EINVAL, "null batch");
  b->timing = enable != 0;
  return IS_OK;
}

int is_batch_kernel_ms(const is_batch* b, double* ms) {
  if (!b || !ms) return fail(IS_EINVAL, "null argument");
  *ms = b->last_ms;
  return IS_OK;
}

int is_batch_status(const is_batch* b, int32_t* status) {
  if (!b || !status) return fail(IS_EINVAL, "null argument");
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  for (int j = 0; j < b->njobs; ++j) status[j] = b->status[j];
  return IS_OK;
}

int64_t is_batch_job_total(const is_batch* b, int32_t job) {
  if (!b || job < 0 || job >= b->njobs) return 0;
  return (int64_t)b->jobs[job].rounds * b->jobs[job].N;
}

int is_batch_job_failure(is_batch* b, int32_t job, is_failure* out) {
  BATCH_JOB_CHECK(b, job);
  if (!out) return fail(IS_EINVAL, "null output");
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const BatchJobHost& H = b->jobs[job];
  std::vector<uint64_t> rec((size_t)3 * H.rounds);
  std::vector<int> fit(H.rounds);
  HIP_TRY(locked_memcpy(rec.data(), b->fail.as<uint64_t>() + 3 * H.round_off, rec.size() * 8, hipMemcpyDeviceToHost));
  HIP_TRY(locked_memcpy(fit.data(), b->fit.as<int>() + H.round_off, fit.size() * sizeof(int), hipMemcpyDeviceToHost));
  is_failure f{};
  for (int t = 0; t < H.rounds; ++t) {
    const uint64_t* w = rec.data() + 3 * t;
    if (w[0] != 0) {
      f.round = t + 1;
      f.count = w[0];
      f.words[0] = w[0];
      f.words[1] = w[1];
      f.words[2] = w[2];
      const uint32_t inv_index = (uint32_t)(w[1] >> 32);
      f.first_logical_index = (int64_t)(uint32_t)(0xffffffffu - inv_index + 1u);
      f.reason_bits = (uint16_t)(w[1] & 0xffff);
      if (w[2]) {
        f.draw_first_logical_index = (int64_t)(uint32_t)(0xffffffffu - (uint32_t)(w[2] >> 32) + 1u);
        f.draw_reason_bits = (uint16_t)(w[2] & 0xffff);
      }
      const uint16_t bits = f.reason_bits;
      f.phase = (w[2] != 0 || (bits & kDrawReasons)) ? IS_PHASE_SAMPLING
                : (bits & (kTargetNaN | kTargetPosInf)) ? IS_PHASE_TARGET
                : (bits & kLogweightInvalid)            ? IS_PHASE_WEIGHT
                                                        : IS_PHASE_DENOMINATOR;
      break;
    }
    if (fit[t] != 0) {
      f.round = t + 1;
      f.phase = fit[t] == 1 ? IS_PHASE_MOMENT : IS_PHASE_FACTORIZATION;
      f.reason_bits = kMomentFitInvalid;
      break;
    }
  }
  *out = f;
  return IS_OK;
}

int is_batch_job_samples(is_batch* b, int32_t job, double* out) {
  BATCH_JOB_CHECK(b, job);
  if (!out) return fail(IS_EINVAL, "null output");
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const BatchJobHost& H = b->jobs[job];
  const int64_t n = (int64_t)H.rounds * H.N;
  const int64_t blocks = (n + 63) / 64;
  std::vector<double> raw((size_t)blocks * 64 * H.P);
  int rc = b->precision == IS_FP32 ? copy_job_T<float>(b, b->X, H.x_off, (int64_t)raw.size(), raw.data())
                                   : copy_job_T<double>(b, b->X, H.x_off, (int64_t)raw.size(), raw.data());
  if (rc) return rc;
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < H.d; ++c) out[i * H.d + c] = raw[(size_t)(bx_base(i, H.P) + 64 * c)];
  return IS_OK;
}

int is_batch_job_logweights(is_batch* b, int32_t job, double* out) {
  BATCH_JOB_CHECK(b, job);
  if (!out) return fail(IS_EINVAL, "null output");
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const BatchJobHost& H = b->jobs[job];
  const int64_t n = (int64_t)H.rounds * H.N;
  return b->precision == IS_FP32 ? copy_job_T<float>(b, b->lw, H.log_off, n, out)
                                 : copy_job_T<double>(b, b->lw, H.log_off, n, out);
}

int is_batch_job_round_summary(is_batch* b, int32_t job, double* ess, double* lognormalizers) {
  BATCH_JOB_CHECK(b, job);
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const BatchJobHost& H = b->jobs[job];
  if (ess)
    HIP_TRY(locked_memcpy(ess, b->ress.as<double>() + H.round_off, H.rounds * sizeof(double), hipMemcpyDeviceToHost));
  if (lognormalizers)
    HIP_TRY(locked_memcpy(lognormalizers, b->rlogz.as<double>() + H.round_off, H.rounds * sizeof(double),
                      hipMemcpyDeviceToHost));
  return IS_OK;
}

int is_batch_final_summary(is_batch* b, double* ess, double* lognormalizers) {
  if (!b) return fail(IS_EINVAL, "null batch");
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const size_t nround = b->ress.bytes / sizeof(double);
  std::vector<double> e(nround), z(nround);
  HIP_TRY(locked_memcpy(e.data(), b->ress.p, nround * sizeof(double), hipMemcpyDeviceToHost));
  HIP_TRY(locked_memcpy(z.data(), b->rlogz.p, nround * sizeof(double), hipMemcpyDeviceToHost));
  for (int j = 0; j < b->njobs; ++j) {
    const BatchJobHost& H = b->jobs[j];
    if (ess) ess[j] = e[H.round_off + H.rounds - 1];
    if (lognormalizers) lognormalizers[j] = z[H.round_off + H.rounds - 1];
  }
  return IS_OK;
}

int is_batch_job_history(is_batch* b, int32_t job, double* means, double* factors, double* lognormalizers) {
  BATCH_JOB_CHECK(b, job);
  if (!b->have_results) return fail(IS_ESTATE, "the batch has not run");
  const BatchJobHost& H = b->jobs[job];
  std::vector<double> slots((size_t)(H.rounds + 1) * H.SS);
  HIP_TRY(locked_memcpy(slots.data(), b->slots64.as<double>() + H.slot_off, slots.size() * sizeof(double),
                    hipMemcpyDeviceToHost));
  const int d = H.d;
  for (int s = 0; s <= H.rounds; ++s)
    unpack_bslot(slots.data() + (size_t)s * H.SS, d, H.P, means ? means + (size_t)s * d : nullptr,
                 factors ? factors + (size_t)s * d * d : nullptr, lognormalizers ? lognormalizers + s : nullptr);
  return IS_OK;
}

int is_batch_job_proposal(is_batch* b, int32_t job, double* location, double* factor, double* lognormalizer) {
  BATCH_JOB_CHECK(b, job);
  const BatchJobHost& H = b->jobs[job];
  std::vector<double> slot(H.SS);
  HIP_TRY(locked_memcpy(slot.data(), b->com64.as<double>() + H.com_off, slot.size() * sizeof(double),
                    hipMemcpyDeviceToHost));
  unpack_bslot(slot.data(), H.d, H.P, location, factor, lognormalizer);
  return IS_OK;
}

int is_batch_export_normals(is_batch* b, int32_t job, int32_t round, double* out, int64_t count) {
  BATCH_JOB_CHECK(b, job);
  const BatchJobHost& H = b->jobs[job];
  if (!out || round < 0 || round >= H.rounds || count != (int64_t)H.d * H.N)
    return fail(IS_EINVAL, "round export needs d * round_size values");
  HIP_TRY(hipSetDevice(b->ctx->device));
  DevBuf tmp;
  int rc = tmp.alloc((size_t)count * sizeof(double));
  if (rc) return rc;
  const uint64_t word = (b->run_index << 32) | (uint64_t)round;
  hipLaunchKernelGGL(k_export_normals, dim3((unsigned)cdiv(H.N, 256)), dim3(256), 0, b->stream, H.key, word, H.N,
                     H.d, kPurposeNormals, tmp.as<double>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, tmp.p, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  return IS_OK;
}

int is_batch_export_radial(is_batch* b, int32_t job, int32_t round, double* normals, int64_t ncount,
                           double* uniforms, int64_t ucount) {
  BATCH_JOB_CHECK(b, job);
  const BatchJobHost& H = b->jobs[job];
  if (round < 0 || round >= H.rounds) return fail(IS_EINVAL, "bad round");
  const int rs = radial_normal_stride(H.family, H.dof);
  const int us = radial_uniform_stride(H.family, H.dof);
  HIP_TRY(hipSetDevice(b->ctx->device));
  const uint64_t word = (b->run_index << 32) | (uint64_t)round;
  auto one = [&](double* out, int64_t count, int per, uint64_t purpose, bool uniform) -> int {
    if (!out || !per) return IS_OK;
    if (count != (int64_t)per * H.N) return fail(IS_EINVAL, "radial export needs per-sample count * round_size values");
    DevBuf tmp;
    int rc = tmp.alloc((size_t)count * sizeof(double));
    if (rc) return rc;
    if (uniform)
      hipLaunchKernelGGL(k_export_uniforms, dim3((unsigned)cdiv(H.N, 256)), dim3(256), 0, b->stream, H.key, word,
                         H.N, per, purpose, tmp.as<double>());
    else
      hipLaunchKernelGGL(k_export_normals, dim3((unsigned)cdiv(H.N, 256)), dim3(256), 0, b->stream, H.key, word,
                         H.N, per, purpose, tmp.as<double>());
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out, tmp.p, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return IS_OK;
  };
  int rc = one(normals, ncount, rs, kPurposeRadialNormals, false);
  if (rc) return rc;
  return one(uniforms, ucount, us, kPurposeRadialUniforms, true);
}

}  // extern "C"
