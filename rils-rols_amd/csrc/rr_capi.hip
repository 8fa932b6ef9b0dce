// rr_capi.hip -- implementation of the C ABI declared in include/rr_fitness.h.
//
// Host-side responsibilities: validate every program of a batch before it
// reaches the GPU (opcodes, operand ranges, stack discipline), stage the
// batch in pinned memory, move it to the device, split FP64 OLS launches so
// their work slabs fit the configured device-memory budget, launch, and copy
// the per-candidate results back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include <chrono>

#include "../../include/rr_fitness.h"
#include "rr_kernels.h"
#include "rr_kernels_fp32.h"
#include "rr_device.h"

// Work memory of rr_fitness_batch_noalloc: one array per GPU, resident from
// the code object's load on that GPU until the process ends.  Host code must
// never pass `scratch` itself to a kernel or a copy (the name is a host-side
// stand-in); it uses the device address that hipGetSymbolAddress returns for
// the current device (scratch_address below).
static __device__ __align__(256) unsigned char scratch[RR_NOALLOC_SCRATCH_BYTES];

namespace {

thread_local std::string g_err;

// per-device address of `scratch`, resolved once per device with that device
// current (rr_create); null until then
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define RR_HIP(call)                                                                         \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return fail(RR_E_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// Bump allocator over the file-scope scratch array, active on the calling
// thread during rr_fitness_batch_noalloc only: every DevBuf the call needs is
// carved from it instead of allocated.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  bool overflow = false;
};
thread_local Arena* t_arena = nullptr;

// Set while a call is planned behind a pending pipelined call
// (rr_fitness_submit): that call's kernels and copies may still read the
// context's buffers, so a buffer that has to grow first waits for them.
thread_local hipStream_t t_busy_stream = nullptr;
void wait_if_busy() {
  if (t_busy_stream) {
    (void)hipStreamSynchronize(t_busy_stream);
    t_busy_stream = nullptr;
  }
}

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  bool view = false;   // p points into another buffer (staged inputs): never freed here
  int ensure(size_t bytes) {
    if (bytes <= cap) return RR_OK;
    if (view) { p = nullptr; cap = 0; view = false; }
    if (t_arena) {
      const size_t off = (t_arena->used + 255) & ~size_t(255);
      if (off > t_arena->size || bytes > t_arena->size - off) {
        t_arena->overflow = true;
        return fail(RR_E_LIMIT, "the call needs more device memory than the no-alloc scratch array holds");
      }
      p = t_arena->base + off;
      cap = bytes;
      t_arena->used = off + bytes;
      return RR_OK;
    }
    if (p) { wait_if_busy(); (void)hipFree(p); p = nullptr; cap = 0; }
    size_t want = std::max<size_t>(bytes + bytes / 2, 4096);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      e = hipMalloc(&p, bytes);
      if (e != hipSuccess) { p = nullptr; return fail(RR_E_NOMEM, "hipMalloc failed for work buffer"); }
      want = bytes;
    }
    cap = want;
    return RR_OK;
  }
  void release() {
    if (p && !view) (void)hipFree(p);
    p = nullptr; cap = 0; view = false;
  }
  // make this buffer a view of `bytes` at q (releasing what it owned)
  void set_view(void* q, size_t bytes) {
    if (!view) release();
    p = q; cap = bytes; view = true;
  }
};

struct HostBuf {
  void* p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return RR_OK;
    if (p) { wait_if_busy(); (void)hipHostFree(p); p = nullptr; cap = 0; }
    size_t want = std::max<size_t>(bytes + bytes / 2, 4096);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) { p = nullptr; return fail(RR_E_NOMEM, "hipHostMalloc failed"); }
    cap = want;
    return RR_OK;
  }
  void release() { if (p) (void)hipHostFree(p); p = nullptr; cap = 0; }
};

struct Dataset {
  bool live = false;
  int64_t n = 0, ldx = 0;
  int32_t d = 0;
  double* X = nullptr;    // d columns, stride ldx
  double* y = nullptr;
  float* X32 = nullptr;   // FP32 copy for the precision form
  double* one = nullptr;  // ldx ones (the free term's column of the FP64 form)
  double* xnorm = nullptr;   // per feature: sums of squares over rows >= 0, >= 1 (chunked order)
  double ybar = 0, sst = 0, ysum = 0, yy = 0;
};

}  // namespace

// An entry of the FP64 planner's table of unique factors (by program hash);
// it belongs to the slab chunk being planned when gen matches.
struct FacSlot {
  uint64_t h;
  int32_t u, gen;
};

struct rr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // side streams (default priority) for the concurrent FP32 bucket kernels,
  // forked from and joined back into `stream` with events
  hipStream_t side[rr::FP32_BUCKETS - 1] = {};
  hipEvent_t fork = nullptr, join[rr::FP32_BUCKETS - 1] = {};
  std::vector<Dataset> ds;
  // FP64 form over a subset (RR_FP32_CHECKED): a stream of its own, forked
  // from and joined back into `stream`
  hipStream_t side64 = nullptr;
  hipEvent_t fork64 = nullptr, join64 = nullptr;
  DevBuf d_cand, d_prog, d_code, d_K, d_slab, d_order32, d_coefs, d_fit, d_misc, d_pred_e, d_dstab, d_cand_ds;
  DevBuf d_f64_off, d_f64_ord;   // FP64 slab offsets and launched position -> candidate
  DevBuf d_fac_prog, d_fac_cand, d_fac_item;   // FP64 factor-evaluation work list
  DevBuf d_prog_item, d_norm_part;              // per program: first work item; per work item: chunk sums
  DevBuf d_info, d_pos_item, d_score_part;      // FP64 scoring: QR outcome per candidate, work list, chunk sums
  DevBuf d_qst, d_qphys, d_qcold, d_qpart;      // FP64 QR state between its kernels
  DevBuf d_rp_off, d_tsst, d_tslist;            // chunked QR: chunk-factor offsets in the slab, state, work list
  DevBuf d_prog_kept, d_fac_kept, d_ustart, d_users, d_alive;   // shared kept columns of equal factors
  DevBuf d_dead;                 // per candidate: decided by the factor-0 pre-pass (OLS mode)
  DevBuf d_blk_cand, d_blk_chunk, d_chunk_base, d_work, d_work_flags;   // FP32 chunk map and partials
  HostBuf h_stage, h_out;
  DevBuf d_stage;                  // device mirror of h_stage: staged inputs arrive in one copy
  // option "graph": the plain FP64 form's launches of a call are captured
  // into a graph and replayed through an executable graph kept per launch
  // sequence (the kernels in order), updated with the call's arguments
  int64_t use_graph = 0;
  struct GraphSlot {
    std::vector<const void*> key;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
  };
  std::vector<GraphSlot> graphs;
  uint64_t graph_tick = 0;
  int64_t graph_made = 0, graph_updated = 0;   // executable graphs instantiated / reused
  size_t up_lo = SIZE_MAX, up_hi = 0;   // the staged bytes not yet copied
  int64_t slab_budget_mb = 32768;
  int64_t fp64_eval_rows = 2;
  int64_t fp64_eval_form = 1;    // the factor interpreter: 0 stack in registers, 1 stack in LDS
  int64_t fp32_rows = 2;
  int64_t fp32_chunk_rows = 4096;
  int64_t fp32_streams = 0;   // 1: bucket kernels on concurrent side streams
  int64_t host_threads = 4;   // host threads for checking a large batch
  void* scratch = nullptr;    // device address of this device's scratch array
  double last_ms = 0.0;
  // phase timing (option "phase_timing"): event pairs handed out per call
  int64_t phase_timing = 0;
  struct PhaseMark { int phase; hipEvent_t a, b; };
  std::vector<hipEvent_t> ev_pool;
  std::vector<PhaseMark> marks;
  size_t ev_used = 0;
  double phase_ms[RR_PHASES] = {0, 0, 0, 0, 0, 0, 0, 0};
  std::vector<FacSlot> fac_table;   // kept across calls: never cleared, only re-stamped
  int32_t fac_gen = 0;
  // pipelined calls (rr_fitness_submit / rr_fitness_wait): at most two in
  // flight, completed in submission order; each has its own pinned result
  // buffer and events.  h2d_done marks the last submitted call's staged
  // copy: the staging buffer is not rewritten before it has completed.
  struct Pending {
    rr_fit* fit_out = nullptr;
    double* coefs_out = nullptr;
    size_t fb = 0, cb = 0;
    bool ran = false;   // completed inside the submit (a form or option that runs synchronously)
    double ms = 0.0;
    std::vector<int64_t> dup_copy;   // candidates that take another's results (copy_duplicates)
  };
  Pending pend[2];
  int n_pend = 0, pend_head = 0;
  HostBuf h_res[2];
  hipEvent_t pev0[2] = {}, pev1[2] = {}, pdone[2] = {};
  hipEvent_t h2d_done = nullptr;
  bool h2d_pending = false;
};

namespace {

int drain(rr_ctx* ctx);   // completes the pending pipelined calls (below)

int set_device(rr_ctx* ctx) {
  if (!ctx) return fail(RR_E_ARG, "null context");
  RR_HIP(hipSetDevice(ctx->device));
  return RR_OK;
}

// Phase timing: begin_phase records a start event on `stream` and returns
// the mark's index (-1 when timing is off); end_phase records its end.
hipEvent_t pool_event(rr_ctx* ctx) {
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[ctx->ev_used++];
}

int begin_phase(rr_ctx* ctx, int phase, hipStream_t stream) {
  if (!ctx->phase_timing) return -1;
  hipEvent_t a = pool_event(ctx), b = pool_event(ctx);
  if (!a || !b || hipEventRecord(a, stream) != hipSuccess) return -1;
  ctx->marks.push_back(rr_ctx::PhaseMark{phase, a, b});
  return (int)ctx->marks.size() - 1;
}

void end_phase(rr_ctx* ctx, int mark, hipStream_t stream) {
  if (mark >= 0) (void)hipEventRecord(ctx->marks[(size_t)mark].b, stream);
}

// after the call's streams have been synchronized
void collect_phases(rr_ctx* ctx) {
  for (int i = 0; i < RR_PHASE_HOST_CHECK; ++i) ctx->phase_ms[i] = 0.0;
  for (const rr_ctx::PhaseMark& m : ctx->marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) ctx->phase_ms[m.phase] += ms;
  }
  ctx->marks.clear();
  ctx->ev_used = 0;
}

Dataset* get_ds(rr_ctx* ctx, int32_t h) {
  if (h < 0 || h >= (int32_t)ctx->ds.size() || !ctx->ds[h].live) return nullptr;
  return &ctx->ds[h];
}

// Host worker threads for the batch check, kept for the life of the process
// (a thread per call costs tens of microseconds to start).  The pool is
// created once and never destroyed: its detached workers sleep on its
// condition variable until the process ends, and destroying that variable
// under them at exit would block the exit.  run(n, fn) calls
// fn(0) here and fn(1) .. fn(n-1) on the workers, and returns when all are
// done; a call that finds the pool busy (another context checking at the
// same moment) returns false and the caller does the work another way.
class WorkerPool {
 public:
  bool run(int n, const std::function<void(int)>& fn) {
    std::unique_lock<std::mutex> busy(use_, std::try_to_lock);
    if (!busy.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> g(m_);
      while ((int)workers_ + 1 < n) {
        try {
          std::thread(&WorkerPool::loop, this).detach();
        } catch (...) {
          break;
        }
        ++workers_;
      }
      if ((int)workers_ + 1 < n) return false;
      job_ = &fn;
      next_ = 1;
      total_ = n;
      left_ = n - 1;
      ++gen_;
    }
    cv_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> g(m_);
    done_.wait(g, [&] { return left_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  void loop() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> g(m_);
    for (;;) {
      cv_.wait(g, [&] { return gen_ != seen; });
      seen = gen_;
      while (job_ && next_ < total_) {
        const int i = next_++;
        const std::function<void(int)>* f = job_;
        g.unlock();
        (*f)(i);
        g.lock();
        if (--left_ == 0) done_.notify_one();
      }
    }
  }
  std::mutex use_, m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* job_ = nullptr;
  int next_ = 0, total_ = 0, left_ = 0;
  unsigned workers_ = 0;
  uint64_t gen_ = 0;
};
WorkerPool& pool() {
  static WorkerPool* p = new WorkerPool;
  return *p;
}

// Validate one program: opcodes, operands and the stack discipline.
int check_program(const uint32_t* code, int32_t len, int32_t d, int32_t n_consts, int* max_depth) {
  if (len <= 0) return fail(RR_E_ARG, "empty program");
  int sp = 0, depth = 0;
  bool acc = false;
  for (int32_t pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc], op = ins & 0xffu, arg = ins >> 8;
    if (op >= RR_OP_BIN) {
      const uint32_t b = op - RR_OP_BIN, bop = b >> 3, mode = b & 7u;
      if (bop > RR_BOP_MAX || mode > RR_M_CR) return fail(RR_E_ARG, "bad binary opcode");
      if (!acc) return fail(RR_E_ARG, "binary op with empty accumulator");
      if (mode == RR_M_S || mode == RR_M_SR) { if (sp == 0) return fail(RR_E_ARG, "stack underflow"); --sp; }
      if ((mode == RR_M_V || mode == RR_M_VR) && (int64_t)arg >= d) return fail(RR_E_ARG, "variable index out of range");
      if ((mode == RR_M_C || mode == RR_M_CR) && (int64_t)arg >= n_consts) return fail(RR_E_ARG, "constant index out of range");
      continue;
    }
    switch (op) {
      case RR_OP_LDV: if ((int64_t)arg >= d) return fail(RR_E_ARG, "variable index out of range"); acc = true; break;
      case RR_OP_LDC: if ((int64_t)arg >= n_consts) return fail(RR_E_ARG, "constant index out of range"); acc = true; break;
      case RR_OP_PUSH:
        if (!acc) return fail(RR_E_ARG, "push of empty accumulator");
        if (++sp > RR_MAX_STACK) return fail(RR_E_LIMIT, "program needs more than RR_MAX_STACK stack slots");
        depth = std::max(depth, sp);
        acc = false;
        break;
      case RR_OP_SIN: case RR_OP_COS: case RR_OP_LN: case RR_OP_EXP: case RR_OP_SQRT: case RR_OP_SQR:
        if (!acc) return fail(RR_E_ARG, "unary op with empty accumulator");
        break;
      default: return fail(RR_E_ARG, "bad opcode");
    }
  }
  if (sp != 0 || !acc) return fail(RR_E_ARG, "program leaves the stack unbalanced");
  if (max_depth) *max_depth = std::max(*max_depth, depth);
  return RR_OK;
}

// Does a (checked) program read a sample variable?  One that does not is a
// constant column.
bool reads_variable(const uint32_t* code, int32_t len) {
  for (int32_t pc = 0; pc < len; ++pc) {
    const uint32_t op = code[pc] & 0xffu;
    if (op == RR_OP_LDV) return true;
    if (op >= RR_OP_BIN) {
      const uint32_t mode = (op - RR_OP_BIN) & 7u;
      if (mode == RR_M_V || mode == RR_M_VR) return true;
    }
  }
  return false;
}

// Per-opcode tables for the batch check: the class of every opcode value
// (0: invalid) and its rough per-sample cost in FP64 instruction units (a
// transcendental costs tens of instructions), so that the check of a whole
// batch runs without a branch per instruction.
enum : uint8_t { OC_BAD = 0, OC_LDV, OC_LDC, OC_PUSH, OC_UN, OC_BIN_STACK, OC_BIN_VAR, OC_BIN_CONST };
// Per opcode, what program_ok's loop needs, packed in one entry: invalid,
// needs the accumulator, operand is a variable / a constant, keeps / sets the
// accumulator, stack change, evaluation cost.
struct OpInfo {
  uint8_t bad, needacc, var, cst, keep, set;
  int8_t dsp;
  uint8_t pad;
  int32_t cost;
};
struct OpTables {
  uint8_t cls[256];
  uint16_t cost[256];
  OpInfo info[256];
  OpTables() {
    for (int op = 0; op < 256; ++op) {
      cls[op] = OC_BAD;
      cost[op] = 1;
      if (op == (int)RR_OP_LDV) cls[op] = OC_LDV;
      else if (op == (int)RR_OP_LDC) cls[op] = OC_LDC;
      else if (op == (int)RR_OP_PUSH) cls[op] = OC_PUSH;
      else if (op >= (int)RR_OP_SIN && op <= (int)RR_OP_SQR) {
        cls[op] = OC_UN;
        cost[op] = op <= (int)RR_OP_EXP ? 40 : op == (int)RR_OP_SQRT ? 10 : 1;
      } else if (op >= (int)RR_OP_BIN) {
        const int bop = (op - (int)RR_OP_BIN) >> 3, m = (op - (int)RR_OP_BIN) & 7;
        if (bop <= (int)RR_BOP_MAX && m <= (int)RR_M_CR)
          cls[op] = (m == (int)RR_M_S || m == (int)RR_M_SR) ? OC_BIN_STACK
                    : (m == (int)RR_M_V || m == (int)RR_M_VR) ? OC_BIN_VAR : OC_BIN_CONST;
        if (bop == (int)RR_BOP_DIV) cost[op] = 8;
      }
    }
    for (int op = 0; op < 256; ++op) {
      const int k = cls[op];
      OpInfo& o = info[op];
      o.bad = k == OC_BAD;
      o.needacc = k >= OC_PUSH;   // push, unary and binary operators
      o.var = k == OC_LDV || k == OC_BIN_VAR;
      o.cst = k == OC_LDC || k == OC_BIN_CONST;
      o.keep = k != OC_PUSH;
      o.set = k == OC_LDV || k == OC_LDC;
      o.dsp = (int8_t)((k == OC_PUSH) - (k == OC_BIN_STACK));
      o.pad = 0;
      o.cost = cost[op];
    }
  }
};
const OpTables g_ops;

// check_program's accept/reject decision without its messages; adds the
// program's cost to *w.
// HASH: *h gets a hash of the program in which a constant operand counts by
// its value (bits), not its pool index, so equal programs hash equal.
template <bool HASH>
bool program_ok(const uint32_t* code, int32_t len, int32_t d, int32_t n_consts, const double* K, int64_t* w,
                uint64_t* h, int32_t* depth) {
  if (len <= 0) return false;
  // branch-free: one table entry per instruction, the state in integers
  uint32_t bad = 0, acc = 0;
  int32_t sp = 0, maxsp = 0;
  int64_t cw = 0;
  uint64_t hh = 0x243F6A8885A308D3ULL;
  const uint32_t ud = d > 0 ? (uint32_t)d : 0u, uk = n_consts > 0 ? (uint32_t)n_consts : 0u;
  for (int32_t pc = 0; pc < len; ++pc) {
    const uint32_t ins = code[pc], op = ins & 0xffu, arg = ins >> 8;
    const OpInfo o = g_ops.info[op];
    cw += o.cost;
    sp += o.dsp;
    maxsp = sp > maxsp ? sp : maxsp;
    bad |= o.bad | (o.needacc & (acc ^ 1u)) | (uint32_t)((uint32_t)sp > (uint32_t)RR_MAX_STACK) |
           (o.var & (uint32_t)(arg >= ud)) | (o.cst & (uint32_t)(arg >= uk));
    acc = o.keep & (acc | o.set);
    if (HASH) {   // a rotate-xor chain
      uint64_t bits;
      memcpy(&bits, K + ((o.cst & (uint32_t)(arg < uk)) ? arg : 0u), 8);
      const uint64_t v = o.cst ? (op ^ (bits * 0x9E3779B97F4A7C15ULL)) : ins;
      hh = ((hh << 23) | (hh >> 41)) ^ v;
    }
  }
  *w += cw;
  *depth = maxsp > *depth ? maxsp : *depth;
  if (HASH) {   // splitmix64 finalizer: every output bit depends on every input bit
    hh ^= hh >> 30; hh *= 0xbf58476d1ce4e5b9ULL;
    hh ^= hh >> 27; hh *= 0x94d049bb133111ebULL;
    *h = hh ^ (hh >> 31);
  }
  return !bad && sp == 0 && acc;
}

// Structure of a batch and every program, each against the feature count of
// its candidate's dataset (cand_d[c], or d for all when cand_d is null).
// prog_cost[p], when asked for, gets program p's per-sample evaluation work
// (rough FP64 instruction units): used only to launch the long evaluations
// first, so that one of them does not start last and end the launch alone;
// it changes no result.  prog_hash[p], when asked for, gets the hash of
// program_ok<true> (equal programs: equal hashes).  threads: host threads it
// may use for a large batch.
int check_batch(const rr_batch* b, int32_t mode, int32_t d, const int32_t* cand_d, std::vector<int32_t>* prog_cost,
                std::vector<uint64_t>* prog_hash = nullptr, int threads = 1, int32_t* max_depth = nullptr) {
  if (!b || b->n_cands < 0) return fail(RR_E_ARG, "bad batch");
  if (b->n_cands == 0) return RR_OK;
  if (!b->cand_start || !b->prog_start || (b->n_consts > 0 && !b->consts)) return fail(RR_E_ARG, "null batch array");
  if (b->cand_start[0] != 0) return fail(RR_E_ARG, "cand_start[0] must be 0");
  for (int32_t c = 0; c < b->n_cands; ++c) {
    const int32_t np = b->cand_start[c + 1] - b->cand_start[c];
    if (np < 0) return fail(RR_E_ARG, "cand_start not monotone");
    if (mode == RR_MODE_RAW && np != 1) return fail(RR_E_ARG, "RAW mode takes exactly one program per candidate");
    if (mode == RR_MODE_OLS && np + 1 > RR_MAX_COLS) return fail(RR_E_LIMIT, "candidate has more than RR_MAX_COLS OLS columns");
  }
  const int32_t n_progs = b->cand_start[b->n_cands];
  if (b->prog_start[0] != 0) return fail(RR_E_ARG, "prog_start[0] must be 0");
  if (b->prog_start[n_progs] > 0 && !b->code) return fail(RR_E_ARG, "null code array");
  for (int32_t p = 0; p < n_progs; ++p)
    if (b->prog_start[p + 1] < b->prog_start[p]) return fail(RR_E_ARG, "prog_start not monotone");
  if (prog_cost) prog_cost->assign(std::max(n_progs, 1), 0);
  if (prog_hash) prog_hash->resize(std::max(n_progs, 1));
  static const double zero = 0.0;   // constant reads of a batch without constants
  const double* K = b->n_consts > 0 ? b->consts : &zero;
  // candidates [c0, c1): the first invalid program, or -1
  // (depth: written once at the end -- the threads' slots share a cache line)
  auto check_range = [&](int32_t c0, int32_t c1, int32_t* bad_prog, int32_t* bad_cand, int32_t* depth_out) {
    int32_t depth_l = 0;
    struct Put { int32_t* o; int32_t* v; ~Put() { *o = *v; } } put{depth_out, &depth_l};
    int32_t* const depth = &depth_l;
    for (int32_t c = c0; c < c1; ++c) {
      const int32_t dc = cand_d ? cand_d[c] : d;
      for (int32_t p = b->cand_start[c]; p < b->cand_start[c + 1]; ++p) {
        const int32_t len = b->prog_start[p + 1] - b->prog_start[p];
        int64_t w = 0;
        const bool ok = prog_hash ? program_ok<true>(b->code + b->prog_start[p], len, dc, b->n_consts, K, &w,
                                                     &(*prog_hash)[p], depth)
                                  : program_ok<false>(b->code + b->prog_start[p], len, dc, b->n_consts, K, &w, nullptr,
                                                      depth);
        if (!ok) { *bad_prog = p; *bad_cand = c; return; }
        if (prog_cost) (*prog_cost)[p] = (int32_t)std::min<int64_t>(w, INT32_MAX);
      }
    }
  };
  // large batches: contiguous candidate ranges of about equal instruction
  // count on host threads; the first invalid program (in batch order) is the
  // one reported, whatever the split
  const int64_t n_ins = b->prog_start[n_progs];
  const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(threads, n_ins / 16384));
  std::vector<int32_t> bad_prog(nt, -1), bad_cand(nt, -1), depth(nt, 0);
  if (nt == 1) {
    check_range(0, b->n_cands, &bad_prog[0], &bad_cand[0], &depth[0]);
  } else {
    std::vector<int32_t> cut(nt + 1, b->n_cands);
    cut[0] = 0;
    for (int t = 1; t < nt; ++t) {   // first candidate whose programs start at or after t/nt of the code
      const int64_t target = n_ins * t / nt;
      int32_t lo = cut[t - 1], hi = b->n_cands;
      while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (b->prog_start[b->cand_start[mid]] < target) lo = mid + 1; else hi = mid;
      }
      cut[t] = lo;
    }
    const std::function<void(int)> part = [&](int t) {
      check_range(cut[t], cut[t + 1], &bad_prog[t], &bad_cand[t], &depth[t]);
    };
    if (!pool().run(nt, part)) {   // the pool is busy: threads of this call's own
      std::vector<std::thread> th;
      th.reserve(nt - 1);
      for (int t = 1; t < nt; ++t) {
        try {
          th.emplace_back(part, t);
        } catch (...) {   // no thread to be had: this range here
          part(t);
        }
      }
      part(0);
      for (auto& x : th) x.join();
    }
  }
  if (max_depth) *max_depth = *std::max_element(depth.begin(), depth.end());
  for (int t = 0; t < nt; ++t) {
    if (bad_prog[t] < 0) continue;
    const int32_t p = bad_prog[t], c = bad_cand[t];
    const int32_t len = b->prog_start[p + 1] - b->prog_start[p];
    int rc = check_program(b->code + b->prog_start[p], len, cand_d ? cand_d[c] : d, b->n_consts, nullptr);   // the message
    if (rc == RR_OK) rc = fail(RR_E_ARG, "invalid program");
    char m[64];
    snprintf(m, sizeof m, " (program %d)", p);
    g_err += m;
    return rc;
  }
  return RR_OK;
}

// Puts `bytes` from host `src` into device buffer `dst`.  Through pinned
// staging: the bytes go to h_stage at stage_off now, dst becomes a view of
// the same offset in the device mirror d_stage, and flush_uploads copies
// everything staged since the last flush in one transfer -- it must run
// before the first kernel that reads dst.  No-alloc calls (no staging): a
// copy of its own, from the source itself.
int upload(rr_ctx* ctx, DevBuf& dst, const void* src, size_t bytes, size_t stage_off) {
  if (t_arena) {
    if (int rc = dst.ensure(bytes)) return rc;
    if (bytes == 0) return RR_OK;
    RR_HIP(hipMemcpyAsync(dst.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    RR_HIP(hipStreamSynchronize(ctx->stream));
    return RR_OK;
  }
  if (stage_off + bytes > ctx->h_stage.cap) return fail(RR_E_ARG, "internal: staging offset past the staging buffer");
  if (int rc = ctx->d_stage.ensure(ctx->h_stage.cap)) return rc;
  if (bytes > 0) memcpy((char*)ctx->h_stage.p + stage_off, src, bytes);
  dst.set_view((char*)ctx->d_stage.p + stage_off, bytes);
  if (bytes > 0) {
    ctx->up_lo = std::min(ctx->up_lo, stage_off);
    ctx->up_hi = std::max(ctx->up_hi, stage_off + bytes);
  }
  return RR_OK;
}

int flush_uploads(rr_ctx* ctx) {
  if (ctx->up_hi > ctx->up_lo) {
    RR_HIP(hipMemcpyAsync((char*)ctx->d_stage.p + ctx->up_lo, (char*)ctx->h_stage.p + ctx->up_lo,
                          ctx->up_hi - ctx->up_lo, hipMemcpyHostToDevice, ctx->stream));
  }
  ctx->up_lo = SIZE_MAX;
  ctx->up_hi = 0;
  return RR_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" {

const char* rr_last_error(void) { return g_err.c_str(); }

int rr_check_batch(const rr_batch* b, int32_t mode, int32_t n_features) {
  if (mode != RR_MODE_OLS && mode != RR_MODE_RAW) return fail(RR_E_ARG, "bad mode");
  if (n_features < 0) return fail(RR_E_ARG, "bad feature count");
  return check_batch(b, mode, n_features, nullptr, nullptr, nullptr, 4);
}

int rr_create(int device, rr_ctx** out) {
  if (!out) return fail(RR_E_ARG, "null output pointer");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(RR_E_HIP, "no HIP device available");
  if (device < 0 || device >= count) return fail(RR_E_ARG, "device index out of range");
  RR_HIP(hipSetDevice(device));
  void* scr = nullptr;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if ((int)g_scratch_ptr.size() < count) g_scratch_ptr.resize(count, nullptr);
    if (!g_scratch_ptr[device]) {
      void* ptr = nullptr;
      const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (se != hipSuccess || !ptr)
        return fail(RR_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
      g_scratch_ptr[device] = ptr;
    }
    scr = g_scratch_ptr[device];
  }
  rr_ctx* c = new rr_ctx;
  c->device = device;
  c->scratch = scr;
  bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
            hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess &&
            hipStreamCreateWithFlags(&c->side64, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&c->fork64, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&c->join64, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; ok && i < rr::FP32_BUCKETS - 1; ++i)
    ok = hipStreamCreateWithFlags(&c->side[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming) == hipSuccess;
  for (int i = 0; ok && i < 2; ++i)
    ok = hipEventCreate(&c->pev0[i]) == hipSuccess && hipEventCreate(&c->pev1[i]) == hipSuccess &&
         hipEventCreateWithFlags(&c->pdone[i], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipEventCreateWithFlags(&c->h2d_done, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    delete c;
    return fail(RR_E_HIP, "stream/event creation failed");
  }
  *out = c;
  return RR_OK;
}

int rr_destroy(rr_ctx* ctx) {
  if (!ctx) return RR_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->n_pend = 0;   // pending pipelined calls are dropped: their outputs are not written
  for (Dataset& d : ctx->ds) {
    if (!d.live) continue;
    (void)hipFree(d.X); (void)hipFree(d.y); (void)hipFree(d.X32); (void)hipFree(d.one); (void)hipFree(d.xnorm);
  }
  DevBuf* bufs[] = {&ctx->d_cand, &ctx->d_prog, &ctx->d_code, &ctx->d_K, &ctx->d_slab, &ctx->d_order32,
                    &ctx->d_coefs, &ctx->d_fit, &ctx->d_misc, &ctx->d_pred_e, &ctx->d_dstab, &ctx->d_cand_ds,
                    &ctx->d_blk_cand, &ctx->d_blk_chunk, &ctx->d_chunk_base, &ctx->d_work, &ctx->d_work_flags,
                    &ctx->d_f64_off, &ctx->d_f64_ord, &ctx->d_dead, &ctx->d_fac_prog, &ctx->d_fac_cand,
                    &ctx->d_fac_item, &ctx->d_prog_item, &ctx->d_norm_part, &ctx->d_info, &ctx->d_pos_item,
                    &ctx->d_score_part, &ctx->d_qst, &ctx->d_qphys, &ctx->d_qcold, &ctx->d_qpart,
                    &ctx->d_prog_kept, &ctx->d_fac_kept, &ctx->d_ustart, &ctx->d_users, &ctx->d_alive,
                    &ctx->d_rp_off, &ctx->d_tsst, &ctx->d_tslist};
  for (DevBuf* b : bufs) b->release();
  ctx->h_stage.release();
  ctx->h_out.release();
  for (HostBuf& h : ctx->h_res) h.release();
  ctx->d_stage.release();
  for (int i = 0; i < 2; ++i) {
    (void)hipEventDestroy(ctx->pev0[i]);
    (void)hipEventDestroy(ctx->pev1[i]);
    (void)hipEventDestroy(ctx->pdone[i]);
  }
  (void)hipEventDestroy(ctx->h2d_done);
  for (rr_ctx::GraphSlot& g : ctx->graphs) if (g.exec) (void)hipGraphExecDestroy(g.exec);
  ctx->graphs.clear();
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipEventDestroy(ctx->fork);
  (void)hipStreamSynchronize(ctx->side64);
  (void)hipStreamDestroy(ctx->side64);
  (void)hipEventDestroy(ctx->fork64);
  (void)hipEventDestroy(ctx->join64);
  for (int i = 0; i < rr::FP32_BUCKETS - 1; ++i) {
    (void)hipStreamSynchronize(ctx->side[i]);
    (void)hipStreamDestroy(ctx->side[i]);
    (void)hipEventDestroy(ctx->join[i]);
  }
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return RR_OK;
}

int rr_dataset_upload(rr_ctx* ctx, const double* X_fm, const double* y, int64_t n, int32_t d, int32_t* handle) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = drain(ctx))) return rc;
  if (!handle || !y || n <= 0 || d < 0 || (d > 0 && !X_fm)) return fail(RR_E_ARG, "bad dataset arguments");
  if (n >= ((int64_t)1 << 40)) return fail(RR_E_ARG, "too many samples");
  Dataset D;
  D.n = n; D.d = d;
  D.ldx = (n + rr::RR_ROW_PAD - 1) / rr::RR_ROW_PAD * rr::RR_ROW_PAD;   // whole tiles, never read past a column
  const size_t xb = (size_t)D.ldx * (size_t)std::max(d, 1) * sizeof(double);
  if (hipMalloc(&D.X, xb) != hipSuccess) return fail(RR_E_NOMEM, "hipMalloc for X failed");
  if (hipMalloc(&D.y, (size_t)D.ldx * sizeof(double)) != hipSuccess) { (void)hipFree(D.X); return fail(RR_E_NOMEM, "hipMalloc for y failed"); }
  if (hipMalloc(&D.X32, (size_t)D.ldx * std::max(d, 1) * sizeof(float)) != hipSuccess) {
    (void)hipFree(D.X); (void)hipFree(D.y);
    return fail(RR_E_NOMEM, "hipMalloc for the FP32 copy failed");
  }
  if (hipMalloc(&D.one, (size_t)D.ldx * sizeof(double)) != hipSuccess) {
    (void)hipFree(D.X); (void)hipFree(D.y); (void)hipFree(D.X32);
    return fail(RR_E_NOMEM, "hipMalloc for the column of ones failed");
  }
  RR_HIP(rr::launch_fill_f64(D.one, D.ldx, 1.0, ctx->stream));
  if (hipMalloc(&D.xnorm, (size_t)std::max(d, 1) * 2 * sizeof(double)) != hipSuccess) {
    (void)hipFree(D.X); (void)hipFree(D.y); (void)hipFree(D.X32); (void)hipFree(D.one);
    return fail(RR_E_NOMEM, "hipMalloc for the feature norms failed");
  }
  RR_HIP(hipMemsetAsync(D.X, 0, xb, ctx->stream));
  RR_HIP(hipMemsetAsync(D.y, 0, (size_t)D.ldx * sizeof(double), ctx->stream));
  if (d > 0)
    RR_HIP(hipMemcpy2DAsync(D.X, (size_t)D.ldx * sizeof(double), X_fm, (size_t)n * sizeof(double),
                            (size_t)n * sizeof(double), (size_t)d, hipMemcpyHostToDevice, ctx->stream));
  RR_HIP(hipMemcpyAsync(D.y, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  RR_HIP(rr::launch_to_fp32(D.X, D.X32, (int64_t)D.ldx * std::max(d, 1), ctx->stream));
  RR_HIP(rr::launch_xnorm_fp64(D.X, n, D.ldx, d, D.xnorm, ctx->stream));
  if (ctx->d_misc.ensure(4 * sizeof(double)) != RR_OK) return RR_E_NOMEM;
  RR_HIP(rr::launch_stats_fp64(D.y, n, (double*)ctx->d_misc.p, ctx->stream));
  double st[4];
  RR_HIP(hipMemcpyAsync(st, ctx->d_misc.p, sizeof st, hipMemcpyDeviceToHost, ctx->stream));
  RR_HIP(hipStreamSynchronize(ctx->stream));
  D.ybar = st[0];
  D.sst = st[1];
  D.ysum = st[2];
  D.yy = st[3];
  D.live = true;
  ctx->ds.push_back(D);
  *handle = (int32_t)ctx->ds.size() - 1;
  return RR_OK;
}

int rr_dataset_free(rr_ctx* ctx, int32_t handle) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = drain(ctx))) return rc;
  Dataset* D = get_ds(ctx, handle);
  if (!D) return fail(RR_E_ARG, "bad dataset handle");
  RR_HIP(hipStreamSynchronize(ctx->stream));
  (void)hipFree(D->X); (void)hipFree(D->y); (void)hipFree(D->X32); (void)hipFree(D->one); (void)hipFree(D->xnorm);
  *D = Dataset();
  return RR_OK;
}

int rr_dataset_stats(rr_ctx* ctx, int32_t handle, double* ybar, double* sst) {
  if (!ctx) return fail(RR_E_ARG, "null context");
  Dataset* D = get_ds(ctx, handle);
  if (!D) return fail(RR_E_ARG, "bad dataset handle");
  if (ybar) *ybar = D->ybar;
  if (sst) *sst = D->sst;
  return RR_OK;
}

}  // extern "C"

namespace {

// One batched fitness call.  handles == nullptr: every candidate against
// dataset `handle`; otherwise candidate c against dataset handles[c].
// Capture what `issue` launches on the context's stream into a graph, find
// or make the executable graph for that launch sequence, and launch it
// between ev0 and ev1.  At most GRAPH_SLOTS executable graphs are kept per
// context (least recently used goes).
constexpr size_t GRAPH_SLOTS = 16;
template <class F>
int run_graph(rr_ctx* ctx, F issue) {
  RR_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc_issue = issue();
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(ctx->stream, &g);
  if (rc_issue != RR_OK || e_end != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc_issue != RR_OK ? rc_issue : fail(RR_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e_end));
  }
  // the launch sequence: the kernel of every node, in capture order
  size_t nn = 0;
  int rc = RR_OK;
  std::vector<const void*> key;
  if (hipGraphGetNodes(g, nullptr, &nn) == hipSuccess) {
    std::vector<hipGraphNode_t> nodes(nn);
    if (nn == 0 || hipGraphGetNodes(g, nodes.data(), &nn) == hipSuccess) {
      key.reserve(nn);
      for (hipGraphNode_t nd : nodes) {
        hipGraphNodeType ty;
        hipKernelNodeParams kp{};
        if (hipGraphNodeGetType(nd, &ty) == hipSuccess && ty == hipGraphNodeTypeKernel &&
            hipGraphKernelNodeGetParams(nd, &kp) == hipSuccess)
          key.push_back(kp.func);
        else
          key.push_back((const void*)(uintptr_t)(0x1000 + (int)ty));
      }
    }
  }
  rr_ctx::GraphSlot* slot = nullptr;
  for (rr_ctx::GraphSlot& s : ctx->graphs)
    if (s.exec && s.key == key) { slot = &s; break; }
  bool ready = false;
  if (slot) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    ready = hipGraphExecUpdate(slot->exec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
    if (ready) ++ctx->graph_updated;
    else { (void)hipGraphExecDestroy(slot->exec); slot->exec = nullptr; }
  } else {
    if (ctx->graphs.size() < GRAPH_SLOTS) {
      ctx->graphs.emplace_back();
      slot = &ctx->graphs.back();
    } else {
      slot = &ctx->graphs[0];
      for (rr_ctx::GraphSlot& s : ctx->graphs) if (s.used < slot->used) slot = &s;
      if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
      slot->exec = nullptr;
    }
  }
  if (!ready) {
    const hipError_t e = hipGraphInstantiate(&slot->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      slot->exec = nullptr;
      rc = fail(RR_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    } else {
      slot->key = key;
      ++ctx->graph_made;
    }
  }
  (void)hipGraphDestroy(g);
  if (rc) return rc;
  slot->used = ++ctx->graph_tick;
  RR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  RR_HIP(hipGraphLaunch(slot->exec, ctx->stream));
  RR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  return RR_OK;
}

// Two programs of a batch are the same factor: same opcodes and operands,
// constants compared by value (whatever their pool index).
bool same_program(const rr_batch* b, int32_t p1, int32_t p2) {
  const int32_t l1 = b->prog_start[p1 + 1] - b->prog_start[p1], l2 = b->prog_start[p2 + 1] - b->prog_start[p2];
  if (l1 != l2) return false;
  for (int32_t i = 0; i < l1; ++i) {
    const uint32_t a1 = b->code[b->prog_start[p1] + i], a2 = b->code[b->prog_start[p2] + i];
    const uint32_t op = a1 & 0xffu;
    if (op != (a2 & 0xffu)) return false;
    const bool kop = op == RR_OP_LDC ||
                     (op >= RR_OP_BIN && (((op - RR_OP_BIN) & 7u) == RR_M_C || ((op - RR_OP_BIN) & 7u) == RR_M_CR));
    if (kop) {
      if (memcmp(&b->consts[a1 >> 8], &b->consts[a2 >> 8], 8) != 0) return false;
    } else if (a1 != a2) {
      return false;
    }
  }
  return true;
}

// Candidates of a call whose factor programs equal another's were not
// launched: they take that candidate's results.  dup_copy holds, per such
// candidate, 5 words: the candidate, the one it copies, their first
// coefficient offsets, and the coefficient count (the same for both).
void copy_duplicates(const std::vector<int64_t>& dup_copy, rr_fit* fit_out, double* coefs_out) {
  for (size_t i = 0; i + 4 < dup_copy.size(); i += 5) {
    fit_out[dup_copy[i]] = fit_out[dup_copy[i + 1]];
    if (coefs_out) memcpy(coefs_out + dup_copy[i + 2], coefs_out + dup_copy[i + 3], (size_t)dup_copy[i + 4] * 8);
  }
}

// Completes pipelined call `slot` if it has not completed yet: waits for its
// copies and writes its results to the caller's arrays.
int complete_slot(rr_ctx* ctx, int slot) {
  rr_ctx::Pending& P = ctx->pend[slot];
  if (P.ran) return RR_OK;
  P.ran = true;
  RR_HIP(hipEventSynchronize(ctx->pdone[slot]));
  const char* ho = (const char*)ctx->h_res[slot].p;
  memcpy(P.fit_out, ho, P.fb);
  if (P.cb) memcpy(P.coefs_out, ho + ((P.fb + 255) & ~size_t(255)), P.cb);
  copy_duplicates(P.dup_copy, P.fit_out, P.cb ? P.coefs_out : nullptr);
  float ms = 0.f;
  RR_HIP(hipEventElapsedTime(&ms, ctx->pev0[slot], ctx->pev1[slot]));
  P.ms = ms;
  return RR_OK;
}

// rr_fitness_wait: completes the oldest pending call and retires it.
int wait_oldest(rr_ctx* ctx) {
  if (ctx->n_pend == 0) return fail(RR_E_ARG, "no pipelined call is pending");
  const int slot = ctx->pend_head;
  const int rc = complete_slot(ctx, slot);
  ctx->pend_head = (ctx->pend_head + 1) % 2;
  --ctx->n_pend;
  ctx->last_ms = ctx->pend[slot].ms;
  return rc;
}

// Completes every pending pipelined call, before any other use of the
// context's stream, buffers or datasets; the calls stay queued until their
// rr_fitness_wait, which then returns at once.
int drain(rr_ctx* ctx) {
  int rc = RR_OK;
  for (int i = 0; i < ctx->n_pend; ++i) {
    const int r = complete_slot(ctx, (ctx->pend_head + i) % 2);
    if (r && !rc) rc = r;
  }
  ctx->h2d_pending = false;
  return rc;
}

// async: a pipelined call (rr_fitness_submit; plain FP64 form, no phase
// timing, no graph): launched behind any pending call and left pending, its
// results written by wait_oldest.  Otherwise pending calls complete first.
int fitness_impl(rr_ctx* ctx, int32_t handle, const int32_t* handles, const rr_batch* b, int32_t mode,
                 int32_t precision, double* coefs_out, rr_fit* fit_out, bool async = false) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if (!async && ctx->n_pend > 0 && (rc = drain(ctx))) return rc;
  struct BusyGuard { ~BusyGuard() { t_busy_stream = nullptr; } } busy_guard;
  if (async && ctx->n_pend > 0) t_busy_stream = ctx->stream;
  const int slot = (ctx->pend_head + ctx->n_pend) % 2;   // async: the call's result slot
  if (mode != RR_MODE_OLS && mode != RR_MODE_RAW) return fail(RR_E_ARG, "bad mode");
  if (precision != RR_FP64 && precision != RR_FP32_REFINED && precision != RR_FP32_CHECKED)
    return fail(RR_E_ARG, "bad precision");
  if (precision == RR_FP32_CHECKED && mode != RR_MODE_OLS) return fail(RR_E_ARG, "RR_FP32_CHECKED is an OLS-mode form");
  if (!b || b->n_cands < 0) return fail(RR_E_ARG, "bad batch");
  const int32_t nc = b->n_cands;
  using clk = std::chrono::steady_clock;
  const clk::time_point t_enter = clk::now();
  auto ms_since = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };

  // the datasets of this call: table entries and per-candidate indices
  std::vector<int32_t> uniq;          // dataset handles in table order
  std::vector<int32_t> cand_ds;       // per candidate, index into uniq
  int32_t min_d = 0;
  if (!handles) {
    Dataset* D = get_ds(ctx, handle);
    if (!D) return fail(RR_E_ARG, "bad dataset handle");
    uniq.push_back(handle);
    min_d = D->d;
  } else {
    cand_ds.resize(nc);
    min_d = INT32_MAX;
    for (int32_t c = 0; c < nc; ++c) {
      Dataset* D = get_ds(ctx, handles[c]);
      if (!D) return fail(RR_E_ARG, "bad dataset handle for candidate " + std::to_string(c));
      size_t u = 0;
      while (u < uniq.size() && uniq[u] != handles[c]) ++u;   // few datasets per call
      if (u == uniq.size()) uniq.push_back(handles[c]);
      cand_ds[c] = (int32_t)u;
      min_d = std::min(min_d, D->d);
    }
  }
  std::vector<int32_t> pcost;   // per program: evaluation cost (filled by the check)
  std::vector<uint64_t> phash;  // per program: hash by value (filled by the check, OLS mode)
  int32_t batch_depth = 0;      // the deepest operand stack of any program (filled by the check)
  {
    std::vector<int32_t> cand_d;
    if (handles) {
      cand_d.resize(nc);
      for (int32_t c = 0; c < nc; ++c) cand_d[c] = ctx->ds[uniq[cand_ds[c]]].d;
    }
    if ((rc = check_batch(b, mode, min_d, handles ? cand_d.data() : nullptr, &pcost,
                          mode == RR_MODE_OLS ? &phash : nullptr, (int)ctx->host_threads, &batch_depth)))
      return rc;
  }
  if (!fit_out) return fail(RR_E_ARG, "null fit_out");
  if (mode == RR_MODE_OLS && !coefs_out) return fail(RR_E_ARG, "null coefs_out");
  const clk::time_point t_checked = clk::now();
  if (nc == 0) { ctx->last_ms = 0.0; return RR_OK; }
  const int32_t np = b->cand_start[nc];
  const int32_t ncode = b->prog_start[np];
  const int64_t ncoef = (int64_t)np + nc;

  std::vector<rr::DsDesc> table(uniq.size());
  for (size_t u = 0; u < uniq.size(); ++u) {
    const Dataset& D = ctx->ds[uniq[u]];
    table[u] = rr::DsDesc{D.X, D.X32, D.y, D.n, D.ldx, D.sst, D.ysum, D.yy, D.one, D.xnorm};
  }
  auto ds_of = [&](int32_t c) -> const Dataset& { return ctx->ds[uniq[handles ? cand_ds[c] : 0]]; };

  // Which candidates go to which form.  RR_FP32_CHECKED: candidates with a
  // factor that reads no variable (its column is exactly collinear with the
  // free term, so FP64's rank decision depends on rounding) or with more
  // factors than the FP32 kernels take go straight to the FP64 form, on a
  // stream of their own, while the FP32 form scores the rest.
  const bool checked = precision == RR_FP32_CHECKED;
  std::vector<int32_t> list32, list64;
  if (precision == RR_FP32_REFINED) {
    list32.resize(nc);
    for (int32_t c = 0; c < nc; ++c) list32[c] = c;
  } else if (checked) {
    for (int32_t c = 0; c < nc; ++c) {
      const int32_t p0 = b->cand_start[c], p1 = b->cand_start[c + 1];
      bool to64 = p1 - p0 > rr::FP32_BUCKET_FACTORS[rr::FP32_BUCKETS - 2];
      for (int32_t p = p0; p < p1 && !to64; ++p)
        to64 = !reads_variable(b->code + b->prog_start[p], b->prog_start[p + 1] - b->prog_start[p]);
      (to64 ? list64 : list32).push_back(c);
    }
  }

  // FP32 plan over list32: candidates in factor-count buckets (OLS) or as
  // given (RAW), and the chunk workgroups of the accumulation pass in that order
  std::vector<int32_t> order, blk_cand, blk_chunk, chunk_base;
  int counts[rr::FP32_BUCKETS] = {0, 0, 0, 0, 0}, blocks[rr::FP32_BUCKETS] = {0, 0, 0, 0, 0};
  int64_t total_chunks = 0;
  const int32_t n32 = (int32_t)list32.size();
  if (n32 > 0) {
    order.resize(n32);
    if (mode == RR_MODE_RAW) {
      order = list32;
      counts[0] = n32;
    } else {
      std::vector<int8_t> bucket(n32);
      for (int32_t q = 0; q < n32; ++q) {
        const int32_t c = list32[q];
        const int nfc = b->cand_start[c + 1] - b->cand_start[c];
        int kb = 0;
        while (kb < rr::FP32_BUCKETS - 1 && nfc > rr::FP32_BUCKET_FACTORS[kb]) ++kb;
        bucket[q] = (int8_t)kb;
        ++counts[kb];
      }
      int start[rr::FP32_BUCKETS];
      for (int kb = 0, acc = 0; kb < rr::FP32_BUCKETS; ++kb) { start[kb] = acc; acc += counts[kb]; }
      for (int32_t q = 0; q < n32; ++q) order[start[bucket[q]]++] = list32[q];
    }
    chunk_base.assign(nc, 0);
    for (int kb = 0, q = 0; kb < rr::FP32_BUCKETS; ++kb) {
      for (int i = 0; i < counts[kb]; ++i, ++q) {
        const int32_t c = order[q];
        if (kb == rr::FP32_BUCKETS - 1 && mode == RR_MODE_OLS) continue;   // flagged without evaluation
        const int64_t nch = (ds_of(c).n + ctx->fp32_chunk_rows - 1) / ctx->fp32_chunk_rows;
        if (total_chunks + nch > INT32_MAX) return fail(RR_E_LIMIT, "too many FP32 chunks in one call");
        chunk_base[c] = (int32_t)total_chunks;
        for (int64_t ch = 0; ch < nch; ++ch) { blk_cand.push_back(c); blk_chunk.push_back((int32_t)ch); }
        total_chunks += nch;
        blocks[kb] += (int)nch;
      }
    }
  }

  // stage inputs
  const size_t s_cand = align_up((size_t)(nc + 1) * 4), s_prog = align_up((size_t)(np + 1) * 4);
  const size_t s_code = align_up((size_t)ncode * 4), s_K = align_up((size_t)std::max(b->n_consts, 1) * 8);
  const size_t s_off = align_up((size_t)nc * 8);
  const size_t s_tab = align_up(table.size() * sizeof(rr::DsDesc)), s_cds = align_up((size_t)nc * 4);
  const size_t s_chunks = 2 * align_up((size_t)total_chunks * 4) + align_up(chunk_base.size() * 4);
  // FP64 slab offsets, order and the three arrays of the evaluation work list
  const size_t s_f64 = align_up((size_t)nc * 8) + align_up((size_t)nc * 4) + 4 * align_up((size_t)(np + 1) * 4) +
                       align_up((size_t)(nc + 1) * 4) +       // the scoring work list
                       2 * align_up((size_t)(np + 1) * 8) +   // kept-column offsets per program and per unique factor
                       2 * align_up((size_t)(np + 1) * 4) +   // the users of each unique factor
                       align_up((size_t)nc * 8);              // chunk-factor offsets of the chunked QR
  if (ctx->h2d_pending) {   // the pending call's staged copy still reads the staging buffer
    RR_HIP(hipEventSynchronize(ctx->h2d_done));
    ctx->h2d_pending = false;
  }
  if (!t_arena && ctx->h_stage.ensure(s_cand + s_prog + s_code + s_K + s_off + s_tab + s_cds + s_chunks + 2 * s_f64) != RR_OK)
    return RR_E_NOMEM;
  size_t so = 0;
  if ((rc = upload(ctx, ctx->d_cand, b->cand_start, (size_t)(nc + 1) * 4, so))) return rc;
  so += s_cand;
  if ((rc = upload(ctx, ctx->d_prog, b->prog_start, (size_t)(np + 1) * 4, so))) return rc;
  so += s_prog;
  if (ncode > 0) { if ((rc = upload(ctx, ctx->d_code, b->code, (size_t)ncode * 4, so))) return rc; }
  else if (ctx->d_code.ensure(4) != RR_OK) return RR_E_NOMEM;
  so += s_code;
  {
    static const double zero = 0.0;   // a batch without constants: one zero, never read
    const double* kp = b->n_consts > 0 ? b->consts : &zero;
    if ((rc = upload(ctx, ctx->d_K, kp, (size_t)std::max(b->n_consts, 1) * 8, so))) return rc;
  }
  so += s_K;
  if ((rc = upload(ctx, ctx->d_dstab, table.data(), table.size() * sizeof(rr::DsDesc), so))) return rc;
  so += s_tab;
  const int32_t* dcds = nullptr;
  if (handles) {
    if ((rc = upload(ctx, ctx->d_cand_ds, cand_ds.data(), (size_t)nc * 4, so))) return rc;
    dcds = (const int32_t*)ctx->d_cand_ds.p;
  }
  so += s_cds;
  if (ctx->d_fit.ensure((size_t)nc * sizeof(rr_fit)) != RR_OK) return RR_E_NOMEM;
  if (ctx->d_coefs.ensure((size_t)std::max<int64_t>(ncoef, 1) * 8) != RR_OK) return RR_E_NOMEM;
  const int32_t* ddead = nullptr;   // OLS mode: written by the pre-pass for every candidate
  if (mode == RR_MODE_OLS) {
    if (ctx->d_dead.ensure((size_t)std::max(nc, 1) * 4) != RR_OK) return RR_E_NOMEM;
    if (ctx->d_info.ensure((size_t)std::max(nc, 1) * 4) != RR_OK) return RR_E_NOMEM;
    ddead = (const int32_t*)ctx->d_dead.p;
  }

  const int32_t* dc = (const int32_t*)ctx->d_cand.p;
  const int32_t* dp = (const int32_t*)ctx->d_prog.p;
  const uint32_t* dcode = (const uint32_t*)ctx->d_code.p;
  const double* dK = (const double*)ctx->d_K.p;
  const rr::DsDesc* dtab = (const rr::DsDesc*)ctx->d_dstab.p;
  rr_fit* dfit = (rr_fit*)ctx->d_fit.p;
  double* dcoef = (double*)ctx->d_coefs.p;

  // FP32: order, chunk map and workspace
  rr::Fp32Launch L32{};
  if (n32 > 0) {
    if ((rc = upload(ctx, ctx->d_order32, order.data(), (size_t)n32 * 4, so))) return rc;
    so += s_off;
    if ((rc = upload(ctx, ctx->d_blk_cand, blk_cand.data(), (size_t)total_chunks * 4, so))) return rc;
    so += align_up((size_t)total_chunks * 4);
    if ((rc = upload(ctx, ctx->d_blk_chunk, blk_chunk.data(), (size_t)total_chunks * 4, so))) return rc;
    so += align_up((size_t)total_chunks * 4);
    if ((rc = upload(ctx, ctx->d_chunk_base, chunk_base.data(), chunk_base.size() * 4, so))) return rc;
    so += align_up(chunk_base.size() * 4);
    if (ctx->d_work.ensure((size_t)std::max<int64_t>(total_chunks, 1) * rr::FP32_ACC_STRIDE * 8) != RR_OK ||
        ctx->d_work_flags.ensure((size_t)std::max<int64_t>(total_chunks, 1) * 8) != RR_OK)
      return RR_E_NOMEM;
    L32.cand_start = dc; L32.prog_start = dp; L32.code = dcode; L32.K = dK; L32.ds = dtab; L32.cand_ds = dcds;
    L32.coefs = dcoef; L32.fit = dfit; L32.dead = ddead; L32.n_cands = n32; L32.raw = mode == RR_MODE_RAW ? 1 : 0;
    L32.rows = (int)ctx->fp32_rows;
    L32.blk_cand = (const int32_t*)ctx->d_blk_cand.p; L32.blk_chunk = (const int32_t*)ctx->d_blk_chunk.p;
    L32.chunk_base = (const int32_t*)ctx->d_chunk_base.p; L32.chunk_rows = ctx->fp32_chunk_rows;
    L32.part = (double*)ctx->d_work.p; L32.flags = (uint32_t*)ctx->d_work_flags.p;
  } else {
    so += s_off + s_chunks;
  }

  // FP64 OLS launches over a candidate list (null: every candidate), chunked
  // so that each launch's slab fits the budget.  The device buffers are
  // reused by a second call only after the host has synchronized with the
  // first; its staging goes to slot 0 or 1 of the pinned buffer.
  const size_t so64 = so;
  // FP64 OLS launches over a candidate list (null: every candidate), longest
  // first, chunked so that each launch's slab fits the budget.  plan64 does
  // the host work and the uploads, launch64 only launches.
  // positions c0..c1-1 share one slab; their unique factors are
  // p0..p0+npair-1 of the evaluation work list, with nitem work items
  // sitem chunk work items of the QR and scoring (sper each, or 0).  The
  // first ne positions (more than TS_MAX_COLS columns) take the
  // step-synchronous QR: per step s the positions taking part (step_pos[s]:
  // those with more than s columns, a prefix) and their chunk work items
  // (step_items[s]).  The rest take the chunked QR in three column buckets:
  // positions ts_pos[i] .. ts_pos[i+1]-1 and chunk work items ts_it[i] ..
  // ts_it[i+1]-1 have TS_NC[i] columns with y (positions and items relative to
  // the chunk's first).
  struct Chunk64 {
    int c0, c1, maxk, p0, npair, nitem, per, sitem, sper, ne, depth;
    int ts_pos[4], ts_it[4];
    int ts_maxch;       // the most chunk work items of a chunked-QR position
    int64_t ts_lw[4];   // the buckets' work-item descriptors: first word, and the words in all
    std::vector<int32_t> step_pos, step_items;
  };
  static constexpr int TS_NC[3] = {10, 5, 3};
  // a chunked-QR position has as many chunk work items as scoring ones
  static_assert(rr::TS_ROWS == rr::EVAL_CHUNK_ROWS, "chunked QR and scoring share the chunk numbering");
  auto plan64 = [&](const std::vector<int32_t>* list_in, int slot, std::vector<Chunk64>& chunks) -> int {
    chunks.clear();
    const int32_t m = list_in ? (int32_t)list_in->size() : nc;
    std::vector<int32_t> lpt(m);
    {
      // counting sort by cost, longest first (ties in list order): the
      // candidate's programs plus 24 k^2 for the QR passes
      constexpr int64_t NB = 8192;
      auto cost = [&](int32_t c) {
        const int64_t k = b->cand_start[c + 1] - b->cand_start[c] + 1;
        int64_t w = 24 * k * k;
        for (int32_t p = b->cand_start[c]; p < b->cand_start[c + 1]; ++p) w += pcost[p];
        return w;
      };
      auto key = [&](int32_t c) { return (int)(NB - 1 - std::min<int64_t>(cost(c), NB - 1)); };
      std::vector<int32_t> cnt(NB + 1, 0);
      for (int32_t q = 0; q < m; ++q) ++cnt[key(list_in ? (*list_in)[q] : q)];
      for (int64_t i = 0, a = 0; i <= NB; ++i) { const int32_t t = cnt[i]; cnt[i] = (int32_t)a; a += t; }
      for (int32_t q = 0; q < m; ++q) {
        const int32_t c = list_in ? (*list_in)[q] : q;
        lpt[cnt[key(c)]++] = c;
      }
    }
    {
      // then by column count, most first, keeping that order within a count:
      // step p of the QR takes the positions with more than p columns, a prefix
      std::vector<int32_t> cnt(RR_MAX_COLS + 2, 0), tmp(m);
      auto kk = [&](int32_t c) { return RR_MAX_COLS + 1 - (b->cand_start[c + 1] - b->cand_start[c] + 1); };
      for (int32_t q = 0; q < m; ++q) ++cnt[kk(lpt[q])];
      for (int i = 0, acc = 0; i < RR_MAX_COLS + 2; ++i) { const int32_t t = cnt[i]; cnt[i] = acc; acc += t; }
      for (int32_t q = 0; q < m; ++q) tmp[cnt[kk(lpt[q])]++] = lpt[q];
      lpt.swap(tmp);
    }
    if (m == 0) return RR_OK;
    // Slab: per position its k+1 working columns (A, Q^T y), then the kept
    // columns of the factors it is the first in its chunk to use.  Factors
    // that are the same program over the same dataset (same opcodes and
    // operands, constants compared by value) share one kept column and its
    // norms: one evaluation.
    std::vector<int64_t> off(m), rp_off(m, 0);
    std::vector<int64_t> prog_kept(std::max(np, 1), 0);
    std::vector<int32_t> ufac(std::max(np, 1), -1);   // per non-bare program: its unique factor
    std::vector<int32_t> u_prog, u_cand;              // per unique factor: a program and candidate using it
    std::vector<int64_t> u_kept;                      // per unique factor: its kept column
    std::vector<int32_t> use_prog, use_cand;          // every non-bare program in the list, with its candidate
    auto is_const_op = [](uint32_t op) {
      return op == RR_OP_LDC ||
             (op >= RR_OP_BIN && (((op - RR_OP_BIN) & 7u) == RR_M_C || ((op - RR_OP_BIN) & 7u) == RR_M_CR));
    };
    auto dsi_of = [&](int32_t c) { return handles ? cand_ds[c] : 0; };
    auto same_prog = [&](int32_t p1, int32_t p2) {
      const int32_t l1 = b->prog_start[p1 + 1] - b->prog_start[p1], l2 = b->prog_start[p2 + 1] - b->prog_start[p2];
      if (l1 != l2) return false;
      for (int32_t i = 0; i < l1; ++i) {
        const uint32_t a1 = b->code[b->prog_start[p1] + i], a2 = b->code[b->prog_start[p2] + i];
        if ((a1 & 0xffu) != (a2 & 0xffu)) return false;
        if (is_const_op(a1 & 0xffu)) {
          if (memcmp(&b->consts[a1 >> 8], &b->consts[a2 >> 8], 8) != 0) return false;
        } else if (a1 != a2) {
          return false;
        }
      }
      return true;
    };
    auto is_bare = [&](int32_t p) {
      return b->prog_start[p + 1] - b->prog_start[p] == 1 && (b->code[b->prog_start[p]] & 0xffu) == RR_OP_LDV;
    };
    // no-alloc calls: half of what is left of the scratch array (the rest
    // holds the per-candidate and per-chunk buffers carved after the slab)
    const int64_t budget = t_arena ? (int64_t)std::max<size_t>((t_arena->size - std::min(t_arena->used, t_arena->size)) / 2, 1)
                                   : std::max<int64_t>(ctx->slab_budget_mb, 1) << 20;
    int64_t max_words = 0;
    // per chunk, the unique factors by hash: open addressing, linear probing;
    // an entry belongs to the current chunk when its gen matches, so the
    // table is not cleared between chunks or calls, and it holds at least
    // twice the programs of the list
    int64_t nprog = 0;
    for (int32_t q = 0; q < m; ++q) nprog += b->cand_start[lpt[q] + 1] - b->cand_start[lpt[q]];
    size_t tsize = 64;
    while (tsize < (size_t)(2 * nprog)) tsize <<= 1;
    std::vector<FacSlot>& table = ctx->fac_table;
    if (table.size() < tsize || ctx->fac_gen > INT32_MAX - m - 1) {
      table.assign(std::max(tsize, table.size()), FacSlot{0, -1, -1});
      ctx->fac_gen = 0;
    }
    tsize = table.size();
    // per program, in program order (one sequential pass over what the check
    // threads wrote, instead of cache misses in launch order): the hash
    // mixed with the dataset, and whether it is a bare variable
    std::vector<uint64_t> ph(std::max(np, 1));
    std::vector<uint8_t> bare(std::max(np, 1), 0);
    for (int32_t q = 0; q < m; ++q) {
      const int32_t c = list_in ? (*list_in)[q] : q;
      const uint64_t mx = (uint64_t)(uint32_t)dsi_of(c) * 0x9E3779B97F4A7C15ULL;
      for (int32_t p = b->cand_start[c]; p < b->cand_start[c + 1]; ++p) {
        ph[p] = phash[p] ^ mx;
        bare[p] = is_bare(p);
      }
    }
    std::vector<int32_t> pfound(std::max(np, 1));   // phase 1: unique factor found in the chunk, or -1
    std::vector<int32_t> pearlier(std::max(np, 1)); // phase 1: same factor earlier in the candidate, or -1
    int32_t gen = ctx->fac_gen;
    auto lookup = [&](int32_t p, int32_t dsi) {
      for (size_t t = ph[p] & (tsize - 1);; t = (t + 1) & (tsize - 1)) {
        const FacSlot& e = table[t];
        if (e.gen != gen) return -1;
        if (e.h == ph[p] && dsi_of(u_cand[e.u]) == dsi && same_prog(u_prog[e.u], p)) return e.u;
      }
    };
    auto insert = [&](int32_t p, int32_t u) {
      size_t t = ph[p] & (tsize - 1);
      while (table[t].gen == gen) t = (t + 1) & (tsize - 1);
      table[t] = FacSlot{ph[p], u, gen};
    };
    for (int q = 0; q < m;) {
      int64_t used = 0, rused = 0;   // slab words; chunk-factor words of the chunked QR
      int q1 = q, maxk = 0;
      const int u0 = (int)u_prog.size();
      ++gen;
      while (q1 < m) {
        const int32_t c = lpt[q1];
        const int32_t pb = b->cand_start[c], pe = b->cand_start[c + 1];
        const int k = pe - pb + 1;
        const int64_t ldx = ds_of(c).ldx;
        const int32_t dsi = dsi_of(c);
        // the kept columns it adds: factors new to the chunk, once each
        int64_t fresh = 0;
        for (int32_t p = pb; p < pe; ++p) {
          if (bare[p]) continue;
          pfound[p] = lookup(p, dsi);
          pearlier[p] = -1;
          if (pfound[p] >= 0) continue;
          for (int32_t p2 = pb; p2 < p && pearlier[p] < 0; ++p2)
            if (!bare[p2] && pfound[p2] < 0 && pearlier[p2] < 0 && ph[p2] == ph[p] && same_prog(p2, p)) pearlier[p] = p2;
          if (pearlier[p] < 0) ++fresh;
        }
        // working columns (A, Q^T y) for the step-synchronous QR, chunk
        // factors for the chunked one (placed after the chunk's columns)
        const bool ts = k <= rr::TS_MAX_COLS;
        const int64_t nb = rr::ts_bucket(k);
        const int64_t rneed = ts ? (ds_of(c).n + rr::TS_ROWS - 1) / rr::TS_ROWS * nb * (nb - 1) : 0;
        const int64_t need = ((ts ? 0 : (int64_t)(k + 1)) + fresh) * ldx;
        if (q1 > q && (used + rused + need + rneed) * 8 > budget) break;
        off[q1] = used;
        if (!ts) used += (int64_t)(k + 1) * ldx;
        rp_off[q1] = rused;
        rused += rneed;
        for (int32_t p = pb; p < pe; ++p) {
          if (bare[p]) continue;
          int32_t u = pfound[p] >= 0 ? pfound[p] : pearlier[p] >= 0 ? ufac[pearlier[p]] : -1;
          if (u < 0) {
            u = (int32_t)u_prog.size();
            u_prog.push_back(p);
            u_cand.push_back(c);
            u_kept.push_back(used);
            used += ldx;
            insert(p, u);
          }
          ufac[p] = u;
          prog_kept[p] = u_kept[u];
          use_prog.push_back(p);
          use_cand.push_back(c);
        }
        maxk = std::max(maxk, k);
        ++q1;
      }
      Chunk64 ch{};
      ch.c0 = q; ch.c1 = q1; ch.maxk = maxk; ch.p0 = u0; ch.npair = (int)u_prog.size() - u0;
      chunks.push_back(ch);
      for (int r = q; r < q1; ++r) rp_off[r] += used;   // the chunk factors follow the chunk's columns
      max_words = std::max(max_words, used + rused);
      q = q1;
    }
    ctx->fac_gen = gen;
    if (ctx->d_slab.ensure((size_t)std::max<int64_t>(max_words, 1) * 8) != RR_OK) return RR_E_NOMEM;
    // factor-evaluation work list: per chunk its unique factors, longest
    // first (counting sort on the evaluation cost), each with
    // ceil(n / EVAL_CHUNK_ROWS) work items
    const int32_t nu = (int32_t)u_prog.size();
    std::vector<int32_t> pprog(std::max(nu, 1)), pcand(std::max(nu, 1)), pitem(std::max(nu, 1)),
        prog_item(std::max(np, 1), 0), unew(std::max(nu, 1));
    std::vector<int64_t> pkept(std::max(nu, 1));
    int64_t max_items = 0;
    for (Chunk64& ch : chunks) {
      constexpr int NB = 4096;
      std::vector<int32_t> cnt(NB + 1, 0);
      auto key = [&](int32_t u) { return NB - 1 - std::min(pcost[u_prog[u]], NB - 1); };
      for (int u = ch.p0; u < ch.p0 + ch.npair; ++u) ++cnt[key(u)];
      for (int i = 0, a = 0; i <= NB; ++i) { const int32_t t = cnt[i]; cnt[i] = a; a += t; }
      std::vector<int32_t> sorted(ch.npair);
      for (int u = ch.p0; u < ch.p0 + ch.npair; ++u) sorted[cnt[key(u)]++] = u;
      int64_t items = 0, per = -1;   // per: the common item count of the factors, or 0
      for (int i = 0; i < ch.npair; ++i) {
        const int32_t u = sorted[i], w = ch.p0 + i;
        const int64_t nch = (ds_of(u_cand[u]).n + rr::EVAL_CHUNK_ROWS - 1) / rr::EVAL_CHUNK_ROWS;
        pprog[w] = u_prog[u];
        pcand[w] = u_cand[u];
        pkept[w] = u_kept[u];
        pitem[w] = (int32_t)items;
        unew[u] = w;
        items += nch;
        per = (per < 0 || per == nch) ? nch : 0;
      }
      if (items > INT32_MAX) return fail(RR_E_LIMIT, "too many factor-evaluation work items in one launch");
      ch.nitem = (int)items;
      ch.depth = batch_depth;   // bounds the operand stack of the chunk's factor programs
      max_items = std::max(max_items, items);
      ch.per = per > 0 ? (int)per : 0;
    }
    for (int32_t p = 0; p < np; ++p)
      if (ufac[p] >= 0) prog_item[p] = pitem[unew[ufac[p]]];
    // per unique factor (work-list order) the candidates using it, for the
    // liveness check after the factor-0 pre-pass
    const int32_t nuse = (int32_t)use_prog.size();
    std::vector<int32_t> ustart(nu + 1, 0), users(std::max(nuse, 1));
    for (int32_t i = 0; i < nuse; ++i) ++ustart[unew[ufac[use_prog[i]]] + 1];
    for (int32_t u = 0; u < nu; ++u) ustart[u + 1] += ustart[u];
    {
      std::vector<int32_t> fill(ustart.begin(), ustart.end() - 1);
      for (int32_t i = 0; i < nuse; ++i) users[fill[unew[ufac[use_prog[i]]]]++] = use_cand[i];
    }
    const size_t base = so64 + (size_t)slot * s_f64;
    const size_t sp = align_up((size_t)(np + 1) * 4), sp8 = align_up((size_t)(np + 1) * 8);
    size_t o = base;
    if ((rc = upload(ctx, ctx->d_f64_off, off.data(), (size_t)m * 8, o))) return rc;
    o += align_up((size_t)nc * 8);
    if ((rc = upload(ctx, ctx->d_f64_ord, lpt.data(), (size_t)m * 4, o))) return rc;
    o += align_up((size_t)nc * 4);
    if (nu > 0) {
      if ((rc = upload(ctx, ctx->d_fac_prog, pprog.data(), (size_t)nu * 4, o))) return rc;
      if ((rc = upload(ctx, ctx->d_fac_cand, pcand.data(), (size_t)nu * 4, o + sp))) return rc;
      if ((rc = upload(ctx, ctx->d_fac_item, pitem.data(), (size_t)nu * 4, o + 2 * sp))) return rc;
    }
    if ((rc = upload(ctx, ctx->d_prog_item, prog_item.data(), prog_item.size() * 4, o + 3 * sp))) return rc;
    {
      size_t o2 = o + 4 * sp + align_up((size_t)(nc + 1) * 4);   // after pos_item
      if ((rc = upload(ctx, ctx->d_prog_kept, prog_kept.data(), prog_kept.size() * 8, o2))) return rc;
      o2 += sp8;
      if ((rc = upload(ctx, ctx->d_fac_kept, pkept.data(), pkept.size() * 8, o2))) return rc;
      o2 += sp8;
      if ((rc = upload(ctx, ctx->d_ustart, ustart.data(), ustart.size() * 4, o2))) return rc;
      o2 += sp;
      if ((rc = upload(ctx, ctx->d_users, users.data(), users.size() * 4, o2))) return rc;
      o2 += sp;
      if ((rc = upload(ctx, ctx->d_rp_off, rp_off.data(), (size_t)m * 8, o2))) return rc;
      if (ctx->d_alive.ensure((size_t)std::max(nu, 1) * 4) != RR_OK) return RR_E_NOMEM;
    }
    if (ctx->d_norm_part.ensure((size_t)std::max<int64_t>(max_items, 1) * 16) != RR_OK) return RR_E_NOMEM;
    // scoring work list: per position, ceil(n / EVAL_CHUNK_ROWS) chunks, numbered per slab chunk
    std::vector<int32_t> pos_item(m + 1);
    int64_t max_sitems = 0, max_qpos = 0, max_qcol = 0, max_qpart = 0, max_lwords = 0;
    for (Chunk64& ch : chunks) {
      int64_t it = 0, per = -1;
      for (int q = ch.c0; q < ch.c1; ++q) {
        const int64_t nch = (ds_of(lpt[q]).n + rr::EVAL_CHUNK_ROWS - 1) / rr::EVAL_CHUNK_ROWS;
        pos_item[q] = (int32_t)it;
        it += nch;
        per = (per < 0 || per == nch) ? nch : 0;
      }
      if (it > INT32_MAX) return fail(RR_E_LIMIT, "too many scoring work items in one launch");
      ch.sitem = (int)it;
      ch.sper = per > 0 ? (int)per : 0;
      max_sitems = std::max(max_sitems, it);
      const int mc = ch.c1 - ch.c0;
      auto kof = [&](int r) { return b->cand_start[lpt[ch.c0 + r] + 1] - b->cand_start[lpt[ch.c0 + r]] + 1; };
      auto item_at = [&](int r) { return r < mc ? pos_item[ch.c0 + r] : ch.sitem; };
      ch.ne = 0;
      while (ch.ne < mc && kof(ch.ne) > rr::TS_MAX_COLS) ++ch.ne;
      for (int st = 0; st < ch.maxk && ch.ne > 0; ++st) {
        int na = 0;
        while (na < ch.ne && kof(na) > st) ++na;
        if (na == 0) break;
        ch.step_pos.push_back(na);
        ch.step_items.push_back(item_at(na));
      }
      // the chunked positions by column bucket (contiguous: sorted by column count, most first)
      ch.ts_pos[0] = ch.ne;
      for (int i = 0; i < 3; ++i) {
        int r = ch.ts_pos[i];
        while (r < mc && rr::ts_bucket(kof(r)) == TS_NC[i]) ++r;
        ch.ts_pos[i + 1] = r;
      }
      if (ch.ts_pos[3] != mc) return fail(RR_E_ARG, "internal: chunked QR positions out of column order");
      for (int i = 0; i < 4; ++i) ch.ts_it[i] = item_at(ch.ts_pos[i]);
      ch.ts_maxch = 0;
      for (int r = ch.ne; r < mc; ++r) ch.ts_maxch = std::max(ch.ts_maxch, item_at(r + 1) - item_at(r));
      ch.ts_lw[0] = 0;
      for (int i = 0; i < 3; ++i)
        ch.ts_lw[i + 1] = ch.ts_lw[i] + (int64_t)(ch.ts_it[i + 1] - ch.ts_it[i]) * rr::ts_desc_words(TS_NC[i]);
      max_lwords = std::max(max_lwords, ch.ts_lw[3]);
      max_qpos = std::max<int64_t>(max_qpos, mc);
      max_qcol = std::max<int64_t>(max_qcol, (int64_t)ch.ne * ch.maxk);
      max_qpart = std::max<int64_t>(max_qpart, (int64_t)item_at(ch.ne) * 2 * ch.maxk);
    }
    pos_item[m] = 0;
    if ((rc = upload(ctx, ctx->d_pos_item, pos_item.data(), pos_item.size() * 4, o + 4 * sp))) return rc;
    if (ctx->d_score_part.ensure((size_t)std::max<int64_t>(max_sitems, 1) * 8) != RR_OK) return RR_E_NOMEM;
    if (ctx->d_qst.ensure((size_t)std::max<int64_t>(max_qpos, 1) * sizeof(rr::QrState)) != RR_OK ||
        ctx->d_tsst.ensure((size_t)std::max<int64_t>(max_qpos, 1) * sizeof(rr::TsState)) != RR_OK ||
        ctx->d_tslist.ensure((size_t)(max_lwords + 2) * 8) != RR_OK ||
        ctx->d_qphys.ensure((size_t)std::max<int64_t>(max_qcol, 1) * 4) != RR_OK ||
        ctx->d_qcold.ensure((size_t)std::max<int64_t>(max_qcol, 1) * 5 * 8) != RR_OK ||
        ctx->d_qpart.ensure((size_t)std::max<int64_t>(max_qpart, 1) * 8) != RR_OK)
      return RR_E_NOMEM;
    return RR_OK;
  };
  auto launch64 = [&](const std::vector<Chunk64>& chunks, hipStream_t stream) -> int {
    if (chunks.empty()) return RR_OK;
    if ((rc = flush_uploads(ctx))) return rc;   // a plan made after the first flush
    if (stream != ctx->stream) {   // after the work already queued on the context's stream
      RR_HIP(hipEventRecord(ctx->fork64, ctx->stream));
      RR_HIP(hipStreamWaitEvent(stream, ctx->fork64, 0));
    }
    // which unique factors a live candidate uses (after the factor-0 pre-pass)
    const int32_t nu = chunks.back().p0 + chunks.back().npair;
    int pm = begin_phase(ctx, RR_PHASE_EVAL, stream);
    RR_HIP(rr::launch_alive_fp64((const int32_t*)ctx->d_ustart.p, (const int32_t*)ctx->d_users.p, ddead, nu,
                                 (int32_t*)ctx->d_alive.p, stream));
    end_phase(ctx, pm, stream);
    for (const Chunk64& ch : chunks) {
      pm = begin_phase(ctx, RR_PHASE_EVAL, stream);
      if (ch.npair > 0) {
        rr::EvalLaunch E;
        E.prog_start = dp; E.code = dcode; E.K = dK; E.ds = dtab; E.cand_ds = dcds;
        E.slab = (double*)ctx->d_slab.p;
        E.fac_prog = (const int32_t*)ctx->d_fac_prog.p + ch.p0;
        E.fac_cand = (const int32_t*)ctx->d_fac_cand.p + ch.p0;
        E.fac_item = (const int32_t*)ctx->d_fac_item.p + ch.p0;
        E.fac_kept = (const int64_t*)ctx->d_fac_kept.p + ch.p0;
        E.alive = (const int32_t*)ctx->d_alive.p + ch.p0;
        E.n_facs = ch.npair; E.n_items = ch.nitem; E.items_per_fac = ch.per;
        E.norm_part = (double*)ctx->d_norm_part.p; E.rows = (int)ctx->fp64_eval_rows; E.depth = ch.depth;
        E.form = (int)ctx->fp64_eval_form;
        RR_HIP(rr::launch_eval_fp64(E, stream));
      }
      end_phase(ctx, pm, stream);
      rr::QrLaunch Q{};
      Q.cand_start = dc; Q.prog_start = dp; Q.code = dcode; Q.ds = dtab; Q.cand_ds = dcds;
      Q.coefs = dcoef; Q.fit = dfit; Q.dead = ddead; Q.info = (int32_t*)ctx->d_info.p;
      Q.order = (const int32_t*)ctx->d_f64_ord.p + ch.c0; Q.slab = (double*)ctx->d_slab.p;
      Q.slab_off = (const int64_t*)ctx->d_f64_off.p + ch.c0;
      Q.pos_item = (const int32_t*)ctx->d_pos_item.p + ch.c0; Q.items_per_pos = ch.sper; Q.n_pos = ch.c1 - ch.c0;
      Q.norm_part = (const double*)ctx->d_norm_part.p; Q.prog_item = (const int32_t*)ctx->d_prog_item.p;
      Q.prog_kept = (const int64_t*)ctx->d_prog_kept.p;
      Q.st = (rr::QrState*)ctx->d_qst.p; Q.phys = (int32_t*)ctx->d_qphys.p; Q.cold = (double*)ctx->d_qcold.p;
      Q.ks = ch.maxk; Q.part = (double*)ctx->d_qpart.p;
      pm = begin_phase(ctx, RR_PHASE_QR, stream);
      if (ch.ne > 0) {
        Q.n_pos = ch.ne;
        if (ch.step_items[0] == ch.ne) {   // one chunk work item per position
          RR_HIP(rr::launch_qr_all1(Q, stream));
        } else {
          RR_HIP(rr::launch_qr_init(Q, stream));
          for (size_t st = 0; st < ch.step_pos.size(); ++st)
            RR_HIP(rr::launch_qr_step(Q, ch.step_pos[st], ch.step_items[st], stream));
          RR_HIP(rr::launch_qr_solve(Q, stream));
        }
      }
      const int mc = ch.c1 - ch.c0;
      if (ch.ne < mc) {
        rr::TsLaunch T{};
        T.cand_start = dc; T.prog_start = dp; T.code = dcode; T.ds = dtab; T.cand_ds = dcds;
        T.coefs = dcoef; T.fit = dfit; T.dead = ddead; T.info = (int32_t*)ctx->d_info.p;
        T.order = (const int32_t*)ctx->d_f64_ord.p + ch.c0; T.slab = (const double*)ctx->d_slab.p;
        T.norm_part = (const double*)ctx->d_norm_part.p; T.prog_item = (const int32_t*)ctx->d_prog_item.p;
        T.prog_kept = (const int64_t*)ctx->d_prog_kept.p;
        T.rp_off = (const int64_t*)ctx->d_rp_off.p + ch.c0;
        T.st = (rr::TsState*)ctx->d_tsst.p; T.rpart = (double*)ctx->d_slab.p;
        // the work list: its three counters, then the descriptors
        T.cnt = (int32_t*)ctx->d_tslist.p; T.desc = (int64_t*)ctx->d_tslist.p + 2;
        for (int i = 0; i < 3; ++i) T.list_base[i] = ch.ts_lw[i];
        RR_HIP(rr::launch_ts_init(T, ch.ne, mc, stream));
        for (int i = 0; i < 3; ++i) RR_HIP(rr::launch_ts_chunk(T, i, ch.ts_it[i + 1] - ch.ts_it[i], stream));
        RR_HIP(rr::launch_ts_solve(T, ch.ne, mc, ch.ts_maxch, stream));
      }
      end_phase(ctx, pm, stream);
      rr::ScoreLaunch Sc{};
      Sc.cand_start = dc; Sc.prog_start = dp; Sc.code = dcode; Sc.K = dK; Sc.ds = dtab; Sc.cand_ds = dcds;
      Sc.coefs = dcoef; Sc.slab = (const double*)ctx->d_slab.p; Sc.prog_kept = (const int64_t*)ctx->d_prog_kept.p;
      Sc.order = (const int32_t*)ctx->d_f64_ord.p + ch.c0; Sc.dead = ddead; Sc.info = (const int32_t*)ctx->d_info.p;
      Sc.pos_item = (const int32_t*)ctx->d_pos_item.p + ch.c0; Sc.n_pos = ch.c1 - ch.c0; Sc.n_items = ch.sitem;
      Sc.items_per_pos = ch.sper; Sc.raw = 0; Sc.rows = (int)ctx->fp64_eval_rows;
      Sc.part = (double*)ctx->d_score_part.p; Sc.fit = dfit;
      pm = begin_phase(ctx, RR_PHASE_SCORE, stream);
      RR_HIP(rr::launch_score_fp64(Sc, stream));
      end_phase(ctx, pm, stream);
    }
    return RR_OK;
  };

  // FP32 launches on the context's stream (bucket kernels back to back, or
  // with option fp32_streams bucket 0 there and the others on side streams)
  auto fp32_run = [&]() -> int {
    if (n32 == 0) return RR_OK;
    const int pm = begin_phase(ctx, RR_PHASE_FP32, ctx->stream);
    const bool side = ctx->fp32_streams != 0 && mode == RR_MODE_OLS;
    hipStream_t st[rr::FP32_BUCKETS];
    st[0] = ctx->stream;
    for (int i = 1; i < rr::FP32_BUCKETS; ++i) st[i] = side ? ctx->side[i - 1] : ctx->stream;
    if (side) {
      RR_HIP(hipEventRecord(ctx->fork, ctx->stream));
      for (int i = 0; i < rr::FP32_BUCKETS - 1; ++i) RR_HIP(hipStreamWaitEvent(ctx->side[i], ctx->fork, 0));
    }
    RR_HIP(rr::launch_fp32(L32, (const int32_t*)ctx->d_order32.p, counts, blocks, st));
    if (side) {
      for (int i = 0; i < rr::FP32_BUCKETS - 1; ++i) {
        RR_HIP(hipEventRecord(ctx->join[i], ctx->side[i]));
        RR_HIP(hipStreamWaitEvent(ctx->stream, ctx->join[i], 0));
      }
    }
    end_phase(ctx, pm, ctx->stream);
    return RR_OK;
  };

  std::vector<int32_t> rescored;   // candidates whose result came from the FP64 form (checked mode)
  std::vector<Chunk64> plan;
  // Candidates with the same factor programs in the same order over the same
  // dataset have the same coefficients and fitness (a local-search sweep
  // holds many: perturbations that only change constants the OLS absorbs).
  // Only the first of each is launched; the others copy its results.
  std::vector<int64_t> dup_copy;
  if (mode == RR_MODE_OLS && precision == RR_FP64) {
    std::vector<int32_t> reps;
    size_t tsz = 64;
    while (tsz < (size_t)2 * (size_t)nc) tsz <<= 1;
    std::vector<int32_t> slot_of(tsz, -1);
    std::vector<uint64_t> ch((size_t)nc);
    for (int32_t c = 0; c < nc; ++c) {
      uint64_t h = 0x9E3779B97F4A7C15ULL * (uint64_t)(uint32_t)((handles ? cand_ds[c] : 0) + 1);
      for (int32_t p = b->cand_start[c]; p < b->cand_start[c + 1]; ++p)
        h = (h ^ phash[p]) * 0xff51afd7ed558ccdULL + (uint64_t)(p - b->cand_start[c]);
      ch[(size_t)c] = h ^ (h >> 29);
      int32_t same = -1;
      size_t t = ch[(size_t)c] & (tsz - 1);
      for (; slot_of[t] >= 0; t = (t + 1) & (tsz - 1)) {
        const int32_t r = slot_of[t];
        if (ch[(size_t)r] != ch[(size_t)c]) continue;
        const int32_t nr = b->cand_start[r + 1] - b->cand_start[r], ncn = b->cand_start[c + 1] - b->cand_start[c];
        if (nr != ncn || (handles && cand_ds[r] != cand_ds[c])) continue;
        bool eq = true;
        for (int32_t i = 0; i < ncn && eq; ++i) eq = same_program(b, b->cand_start[r] + i, b->cand_start[c] + i);
        if (eq) { same = r; break; }
      }
      if (same >= 0) {
        dup_copy.insert(dup_copy.end(), {(int64_t)c, (int64_t)same, (int64_t)b->cand_start[c] + c,
                                         (int64_t)b->cand_start[same] + same,
                                         (int64_t)(b->cand_start[c + 1] - b->cand_start[c]) + 1});
      } else {
        slot_of[t] = c;
        reps.push_back(c);
      }
    }
    if ((rc = plan64(dup_copy.empty() ? nullptr : &reps, 0, plan))) return rc;
  }
  rr::ScoreLaunch raw64{};   // RAW mode, plain FP64: every candidate scored in chunks
  if (mode == RR_MODE_RAW && precision == RR_FP64) {
    std::vector<int32_t> pos_item(nc + 1);
    int64_t it = 0, per = -1;
    for (int32_t c = 0; c < nc; ++c) {
      const int64_t nch = (ds_of(c).n + rr::EVAL_CHUNK_ROWS - 1) / rr::EVAL_CHUNK_ROWS;
      pos_item[c] = (int32_t)it;
      it += nch;
      per = (per < 0 || per == nch) ? nch : 0;
    }
    if (it > INT32_MAX) return fail(RR_E_LIMIT, "too many scoring work items in one launch");
    pos_item[nc] = (int32_t)it;
    if ((rc = upload(ctx, ctx->d_pos_item, pos_item.data(), pos_item.size() * 4, so64))) return rc;
    if (ctx->d_score_part.ensure((size_t)std::max<int64_t>(it, 1) * 8) != RR_OK) return RR_E_NOMEM;
    raw64.cand_start = dc; raw64.prog_start = dp; raw64.code = dcode; raw64.K = dK; raw64.ds = dtab;
    raw64.cand_ds = dcds; raw64.prog_kept = nullptr; raw64.order = nullptr; raw64.dead = nullptr; raw64.info = nullptr;
    raw64.pos_item = (const int32_t*)ctx->d_pos_item.p; raw64.n_pos = nc; raw64.n_items = (int)it;
    raw64.items_per_pos = per > 0 ? (int)per : 0; raw64.raw = 1;
    raw64.rows = (int)ctx->fp64_eval_rows;
    raw64.part = (double*)ctx->d_score_part.p; raw64.fit = dfit;
  }
  if (checked && !list64.empty()) { if ((rc = plan64(&list64, 0, plan))) return rc; }
  // the timed region starts with every input uploaded
  ctx->marks.clear();
  ctx->ev_used = 0;
  const clk::time_point t_launch = clk::now();
  if ((rc = flush_uploads(ctx))) return rc;
  if (async) {   // every input of the plain FP64 form is staged by now
    RR_HIP(hipEventRecord(ctx->h2d_done, ctx->stream));
    ctx->h2d_pending = true;
  }
  hipEvent_t const ev0 = async ? ctx->pev0[slot] : ctx->ev0, ev1 = async ? ctx->pev1[slot] : ctx->ev1;
  auto prepass = [&]() -> int {
    if (mode == RR_MODE_OLS && nc > 0) {
      // candidates whose factor 0 is NaN at an early row >= 1: outcome written here, skipped below
      rr::Nan0Launch L0{dc, dp, dcode, dK, dtab, dcds, dcoef, dfit, (int32_t*)ctx->d_dead.p, nc};
      const int pm = begin_phase(ctx, RR_PHASE_PREPASS, ctx->stream);
      RR_HIP(rr::launch_nan0_fp64(L0, ctx->stream));
      end_phase(ctx, pm, ctx->stream);
    }
    return RR_OK;
  };
  // the plain FP64 form: every launch on the context's stream
  auto plain64 = [&]() -> int {
    if ((rc = prepass())) return rc;
    if (mode == RR_MODE_RAW) {
      const int pm = begin_phase(ctx, RR_PHASE_SCORE, ctx->stream);
      RR_HIP(rr::launch_score_fp64(raw64, ctx->stream));
      end_phase(ctx, pm, ctx->stream);
      return RR_OK;
    }
    return launch64(plan, ctx->stream);
  };
  const bool graph = precision == RR_FP64 && ctx->use_graph && !ctx->phase_timing && !t_arena;
  if (graph) {   // records ev0 and ev1 around the graph's launch
    if ((rc = run_graph(ctx, plain64))) return rc;
  } else if (precision == RR_FP64) {
    RR_HIP(hipEventRecord(ev0, ctx->stream));
    if ((rc = plain64())) return rc;
  } else if (!checked) {
    RR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if ((rc = prepass())) return rc;
    if ((rc = fp32_run())) return rc;
  } else {
    // FP64 for the predicted candidates beside the FP32 form, then FP64 for
    // the ones the FP32 form could not certify
    RR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if ((rc = prepass())) return rc;
    if ((rc = launch64(plan, ctx->side64))) return rc;
    if ((rc = fp32_run())) return rc;
    if (!list64.empty()) {
      RR_HIP(hipEventRecord(ctx->join64, ctx->side64));
      RR_HIP(hipStreamWaitEvent(ctx->stream, ctx->join64, 0));
    }
    std::vector<rr_fit> f32(nc);
    RR_HIP(hipMemcpyAsync(f32.data(), dfit, (size_t)nc * sizeof(rr_fit), hipMemcpyDeviceToHost, ctx->stream));
    RR_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> again;
    for (int32_t c : list32) if (f32[c].flags & RR_F_NEEDS64) again.push_back(c);
    if (!again.empty()) {
      if ((rc = plan64(&again, 1, plan))) return rc;
      if ((rc = launch64(plan, ctx->stream))) return rc;
    }
    rescored = list64;
    rescored.insert(rescored.end(), again.begin(), again.end());
  }
  if (!graph) RR_HIP(hipEventRecord(ev1, ctx->stream));

  // results: through the pinned buffer h_out (no-alloc calls: straight into
  // the caller's arrays, no host allocation); a pipelined call's through its
  // slot's buffer, completed by wait_oldest
  const size_t fb = (size_t)nc * sizeof(rr_fit), cb = mode == RR_MODE_OLS ? (size_t)ncoef * 8 : 0;
  if (async) {
    if (ctx->h_res[slot].ensure(align_up(fb) + cb) != RR_OK) return RR_E_NOMEM;
    char* ho = (char*)ctx->h_res[slot].p;
    RR_HIP(hipMemcpyAsync(ho, dfit, fb, hipMemcpyDeviceToHost, ctx->stream));
    if (cb) RR_HIP(hipMemcpyAsync(ho + align_up(fb), dcoef, cb, hipMemcpyDeviceToHost, ctx->stream));
    RR_HIP(hipEventRecord(ctx->pdone[slot], ctx->stream));
    rr_ctx::Pending& P = ctx->pend[slot];
    P.fit_out = fit_out; P.coefs_out = coefs_out; P.fb = fb; P.cb = cb; P.ran = false; P.ms = 0.0;
    P.dup_copy.swap(dup_copy);
    ++ctx->n_pend;
    ctx->phase_ms[RR_PHASE_HOST_CHECK] = ms_since(t_enter, t_checked);
    ctx->phase_ms[RR_PHASE_HOST_PLAN] = ms_since(t_checked, t_launch);
    return RR_OK;
  }
  if (!t_arena && ctx->h_out.ensure(align_up(fb) + cb) == RR_OK) {
    char* ho = (char*)ctx->h_out.p;
    RR_HIP(hipMemcpyAsync(ho, dfit, fb, hipMemcpyDeviceToHost, ctx->stream));
    if (cb) RR_HIP(hipMemcpyAsync(ho + align_up(fb), dcoef, cb, hipMemcpyDeviceToHost, ctx->stream));
    RR_HIP(hipStreamSynchronize(ctx->stream));
    memcpy(fit_out, ho, fb);
    if (cb) memcpy(coefs_out, ho + align_up(fb), cb);
  } else {
    RR_HIP(hipMemcpyAsync(fit_out, dfit, fb, hipMemcpyDeviceToHost, ctx->stream));
    if (cb) RR_HIP(hipMemcpyAsync(coefs_out, dcoef, cb, hipMemcpyDeviceToHost, ctx->stream));
    RR_HIP(hipStreamSynchronize(ctx->stream));
  }
  copy_duplicates(dup_copy, fit_out, cb ? coefs_out : nullptr);
  for (int32_t c : rescored) fit_out[c].flags |= RR_F_FP64;
  float ms = 0.f;
  RR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  const clk::time_point t_done = clk::now();
  if (!ctx->marks.empty()) RR_HIP(hipStreamSynchronize(ctx->side64));
  collect_phases(ctx);
  ctx->phase_ms[RR_PHASE_HOST_CHECK] = ms_since(t_enter, t_checked);
  ctx->phase_ms[RR_PHASE_HOST_PLAN] = ms_since(t_checked, t_launch);
  ctx->phase_ms[RR_PHASE_HOST_TAIL] = std::max(0.0, ms_since(t_launch, t_done) - (double)ms);
  return RR_OK;
}

int submit_impl(rr_ctx* ctx, int32_t handle, const int32_t* handles, const rr_batch* b, int32_t mode,
                double* coefs_out, rr_fit* fit_out) {
  if (!ctx) return fail(RR_E_ARG, "null context");
  if (ctx->n_pend >= 2) return fail(RR_E_ARG, "two calls are pending: complete one with rr_fitness_wait first");
  const bool async = !ctx->phase_timing && !ctx->use_graph && b && b->n_cands > 0;
  if (async) return fitness_impl(ctx, handle, handles, b, mode, RR_FP64, coefs_out, fit_out, true);
  // a call that runs synchronously (empty batch, phase timing or graph replay on): its
  // results are written now, and its wait completes at once
  if (int rc = fitness_impl(ctx, handle, handles, b, mode, RR_FP64, coefs_out, fit_out)) return rc;
  const int slot = (ctx->pend_head + ctx->n_pend) % 2;
  ctx->pend[slot] = rr_ctx::Pending{};
  ctx->pend[slot].ran = true;
  ctx->pend[slot].ms = ctx->last_ms;
  ++ctx->n_pend;
  return RR_OK;
}

}  // namespace

extern "C" {

int rr_fitness_batch(rr_ctx* ctx, int32_t handle, const rr_batch* b, int32_t mode, int32_t precision,
                     double* coefs_out, rr_fit* fit_out) {
  return fitness_impl(ctx, handle, nullptr, b, mode, precision, coefs_out, fit_out);
}

int rr_fitness_submit(rr_ctx* ctx, int32_t handle, const rr_batch* b, int32_t mode, double* coefs_out,
                      rr_fit* fit_out) {
  return submit_impl(ctx, handle, nullptr, b, mode, coefs_out, fit_out);
}

int rr_fitness_submit_multi(rr_ctx* ctx, const int32_t* handles, const rr_batch* b, int32_t mode, double* coefs_out,
                            rr_fit* fit_out) {
  if (b && b->n_cands > 0 && !handles) return fail(RR_E_ARG, "null dataset handles");
  return submit_impl(ctx, -1, handles, b, mode, coefs_out, fit_out);
}

int rr_fitness_wait(rr_ctx* ctx) {
  if (!ctx) return fail(RR_E_ARG, "null context");
  (void)hipSetDevice(ctx->device);
  return wait_oldest(ctx);
}

int rr_fitness_multi(rr_ctx* ctx, const int32_t* handles, const rr_batch* b, int32_t mode, int32_t precision,
                     double* coefs_out, rr_fit* fit_out) {
  if (b && b->n_cands > 0 && !handles) return fail(RR_E_ARG, "null dataset handles");
  if (!b || b->n_cands == 0) {
    if (!ctx) return fail(RR_E_ARG, "null context");
    if (!b) return fail(RR_E_ARG, "bad batch");
    ctx->last_ms = 0.0;
    return RR_OK;
  }
  return fitness_impl(ctx, -1, handles, b, mode, precision, coefs_out, fit_out);
}

int rr_fitness_batch_noalloc(rr_ctx* ctx, int32_t handle, const rr_batch* b, int32_t mode, double* coefs_out,
                             rr_fit* fit_out) {
  if (!ctx) return fail(RR_E_ARG, "null context");
  if (ctx->n_pend > 0) {
    (void)hipSetDevice(ctx->device);
    if (int rc = drain(ctx)) return rc;
  }
  if (!ctx->scratch) return fail(RR_E_HIP, "no scratch array address for this device");
  // The call runs with every work buffer carved from the scratch array; the
  // context's own buffers are set aside and put back afterwards, untouched.
  DevBuf* bufs[] = {&ctx->d_cand, &ctx->d_prog, &ctx->d_code, &ctx->d_K, &ctx->d_slab, &ctx->d_order32,
                    &ctx->d_coefs, &ctx->d_fit, &ctx->d_misc, &ctx->d_pred_e, &ctx->d_dstab, &ctx->d_cand_ds,
                    &ctx->d_blk_cand, &ctx->d_blk_chunk, &ctx->d_chunk_base, &ctx->d_work, &ctx->d_work_flags,
                    &ctx->d_f64_off, &ctx->d_f64_ord, &ctx->d_dead, &ctx->d_fac_prog, &ctx->d_fac_cand,
                    &ctx->d_fac_item, &ctx->d_prog_item, &ctx->d_norm_part, &ctx->d_info, &ctx->d_pos_item,
                    &ctx->d_score_part, &ctx->d_qst, &ctx->d_qphys, &ctx->d_qcold, &ctx->d_qpart,
                    &ctx->d_prog_kept, &ctx->d_fac_kept, &ctx->d_ustart, &ctx->d_users, &ctx->d_alive,
                    &ctx->d_rp_off, &ctx->d_tsst, &ctx->d_tslist};
  constexpr size_t NB = sizeof(bufs) / sizeof(bufs[0]);
  DevBuf saved[NB];
  for (size_t i = 0; i < NB; ++i) { saved[i] = *bufs[i]; *bufs[i] = DevBuf(); }
  const int64_t timing = ctx->phase_timing;
  ctx->phase_timing = 0;   // phase events would be created on first use
  Arena a;
  a.base = (char*)ctx->scratch;
  a.size = (size_t)RR_NOALLOC_SCRATCH_BYTES;
  t_arena = &a;
  int rc = fitness_impl(ctx, handle, nullptr, b, mode, RR_FP64, coefs_out, fit_out);
  t_arena = nullptr;
  ctx->phase_timing = timing;
  for (size_t i = 0; i < NB; ++i) *bufs[i] = saved[i];
  if (rc != RR_OK && a.overflow) {
    char m[160];
    snprintf(m, sizeof m, "rr_fitness_batch_noalloc: the batch needs more than the %d MiB scratch array "
             "(fewer candidates or samples per call)", RR_NOALLOC_SCRATCH_MB);
    g_err = m;
    rc = RR_E_LIMIT;
  }
  return rc;
}

int rr_predict(rr_ctx* ctx, int32_t handle, const uint32_t* code, int32_t len, const double* consts, int32_t n_consts,
               double* out) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = drain(ctx))) return rc;
  Dataset* D = get_ds(ctx, handle);
  if (!D) return fail(RR_E_ARG, "bad dataset handle");
  if (!code || !out || n_consts < 0 || (n_consts > 0 && !consts)) return fail(RR_E_ARG, "bad predict arguments");
  if ((rc = check_program(code, len, D->d, n_consts, nullptr))) return rc;
  const size_t s_code = align_up((size_t)len * 4);
  if (ctx->h_stage.ensure(s_code + align_up((size_t)std::max(n_consts, 1) * 8)) != RR_OK) return RR_E_NOMEM;
  if ((rc = upload(ctx, ctx->d_code, code, (size_t)len * 4, 0))) return rc;
  {
    static const double zero = 0.0;   // a program without constants: one zero, never read
    if ((rc = upload(ctx, ctx->d_K, n_consts > 0 ? consts : &zero, (size_t)std::max(n_consts, 1) * 8, s_code))) return rc;
  }
  if (ctx->d_misc.ensure((size_t)D->n * 8) != RR_OK) return RR_E_NOMEM;
  if ((rc = flush_uploads(ctx))) return rc;
  RR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  RR_HIP(rr::launch_predict_fp64((const uint32_t*)ctx->d_code.p, len, (const double*)ctx->d_K.p, D->X, D->n, D->ldx,
                                 (double*)ctx->d_misc.p, ctx->stream));
  RR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  RR_HIP(hipMemcpyAsync(out, ctx->d_misc.p, (size_t)D->n * 8, hipMemcpyDeviceToHost, ctx->stream));
  RR_HIP(hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  RR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return RR_OK;
}

int rr_predict_fp32(rr_ctx* ctx, int32_t handle, const uint32_t* code, int32_t len, const double* consts,
                    int32_t n_consts, float* out_v, float* out_e) {
  int rc = set_device(ctx);
  if (rc) return rc;
  if ((rc = drain(ctx))) return rc;
  Dataset* D = get_ds(ctx, handle);
  if (!D) return fail(RR_E_ARG, "bad dataset handle");
  if (!code || !out_v || !out_e || n_consts < 0 || (n_consts > 0 && !consts)) return fail(RR_E_ARG, "bad predict arguments");
  if ((rc = check_program(code, len, D->d, n_consts, nullptr))) return rc;
  const size_t s_code = align_up((size_t)len * 4);
  if (ctx->h_stage.ensure(s_code + align_up((size_t)std::max(n_consts, 1) * 8)) != RR_OK) return RR_E_NOMEM;
  if ((rc = upload(ctx, ctx->d_code, code, (size_t)len * 4, 0))) return rc;
  {
    static const double zero = 0.0;   // a program without constants: one zero, never read
    if ((rc = upload(ctx, ctx->d_K, n_consts > 0 ? consts : &zero, (size_t)std::max(n_consts, 1) * 8, s_code))) return rc;
  }
  if (ctx->d_misc.ensure((size_t)D->n * 4) != RR_OK || ctx->d_pred_e.ensure((size_t)D->n * 4) != RR_OK) return RR_E_NOMEM;
  if ((rc = flush_uploads(ctx))) return rc;
  RR_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  RR_HIP(rr::launch_predict_fp32((const uint32_t*)ctx->d_code.p, len, (const double*)ctx->d_K.p, D->X32, D->n, D->ldx,
                                 (float*)ctx->d_misc.p, (float*)ctx->d_pred_e.p, ctx->stream));
  RR_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  RR_HIP(hipMemcpyAsync(out_v, ctx->d_misc.p, (size_t)D->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RR_HIP(hipMemcpyAsync(out_e, ctx->d_pred_e.p, (size_t)D->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  RR_HIP(hipStreamSynchronize(ctx->stream));
  float ms = 0.f;
  RR_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return RR_OK;
}

int rr_autotune(rr_ctx* ctx, int32_t handle, const rr_batch* b, int32_t mode, int32_t precision, int32_t reps,
                double* best_ms) {
  if (!ctx || !b) return fail(RR_E_ARG, "null argument");
  if (reps < 1) reps = 1;
  if (b->n_cands <= 0) return fail(RR_E_ARG, "autotune needs a non-empty batch");
  const int32_t np = b->cand_start[b->n_cands];
  std::vector<double> coefs((size_t)np + (size_t)b->n_cands + 1);
  std::vector<rr_fit> fit((size_t)b->n_cands);
  struct Cfg { int64_t a, b2, c; };
  std::vector<Cfg> cfgs;
  if (precision == RR_FP64) {
    for (int64_t e : {1, 2, 4})
      for (int64_t f : {0, 1}) cfgs.push_back({e, f, 0});
  } else {
    for (int64_t r : {1, 2})
      for (int64_t ch : {2048, 4096, 8192, 16384})
        for (int64_t st : {0, 1}) cfgs.push_back({r, ch, st});
  }
  const Cfg keep = precision == RR_FP64 ? Cfg{ctx->fp64_eval_rows, ctx->fp64_eval_form, 0}
                                         : Cfg{ctx->fp32_rows, ctx->fp32_chunk_rows, ctx->fp32_streams};
  auto apply = [&](const Cfg& c) {
    if (precision == RR_FP64) { ctx->fp64_eval_rows = c.a; ctx->fp64_eval_form = c.b2; }
    else { ctx->fp32_rows = c.a; ctx->fp32_chunk_rows = c.b2; ctx->fp32_streams = c.c; }
  };
  double best = 1e300;
  Cfg win = keep;
  for (const Cfg& c : cfgs) {
    apply(c);
    int rc = rr_fitness_batch(ctx, handle, b, mode, precision, coefs.data(), fit.data());   // warm-up
    if (rc) { apply(keep); return rc; }
    double t = 1e300;
    for (int32_t i = 0; i < reps; ++i) {
      if ((rc = rr_fitness_batch(ctx, handle, b, mode, precision, coefs.data(), fit.data()))) { apply(keep); return rc; }
      t = std::min(t, ctx->last_ms);
    }
    if (t < best) { best = t; win = c; }
  }
  apply(win);
  if (best_ms) *best_ms = best;
  return RR_OK;
}

int rr_set_option(rr_ctx* ctx, const char* name, int64_t value) {
  if (!ctx || !name) return fail(RR_E_ARG, "null argument");
  const std::string s(name);
  if (s == "slab_budget_mb") { if (value < 1) return fail(RR_E_ARG, "slab_budget_mb must be >= 1"); ctx->slab_budget_mb = value; }
  else if (s == "fp64_eval_rows") {
    if (value != 1 && value != 2 && value != 4) return fail(RR_E_ARG, "fp64_eval_rows must be 1, 2 or 4");
    ctx->fp64_eval_rows = value;
  } else if (s == "graph") {
    if (value != 0 && value != 1) return fail(RR_E_ARG, "graph must be 0 or 1");
    ctx->use_graph = value;
  } else if (s == "fp64_eval_form") {
    if (value != 0 && value != 1) return fail(RR_E_ARG, "fp64_eval_form must be 0 or 1");
    ctx->fp64_eval_form = value;
  } else if (s == "fp32_rows") {
    if (value != 1 && value != 2) return fail(RR_E_ARG, "fp32_rows must be 1 or 2");
    ctx->fp32_rows = value;
  } else if (s == "fp32_streams") {
    if (value != 0 && value != 1) return fail(RR_E_ARG, "fp32_streams must be 0 or 1");
    ctx->fp32_streams = value;
  } else if (s == "host_threads") {
    if (value < 1 || value > 64) return fail(RR_E_ARG, "host_threads must be in [1, 64]");
    ctx->host_threads = value;
  } else if (s == "phase_timing") {
    if (value != 0 && value != 1) return fail(RR_E_ARG, "phase_timing must be 0 or 1");
    ctx->phase_timing = value;
  } else if (s == "fp32_chunk_rows") {
    if (value < 512 || value > (int64_t)1 << 30 || value % 512) return fail(RR_E_ARG, "fp32_chunk_rows must be a multiple of 512 in [512, 2^30]");
    ctx->fp32_chunk_rows = value;
  } else return fail(RR_E_ARG, "unknown option " + s);
  return RR_OK;
}

int rr_get_option(rr_ctx* ctx, const char* name, int64_t* value) {
  if (!ctx || !name || !value) return fail(RR_E_ARG, "null argument");
  const std::string s(name);
  if (s == "slab_budget_mb") *value = ctx->slab_budget_mb;
  else if (s == "fp64_eval_rows") *value = ctx->fp64_eval_rows;
  else if (s == "fp64_eval_form") *value = ctx->fp64_eval_form;
  else if (s == "graph") *value = ctx->use_graph;
  else if (s == "graph_made") *value = ctx->graph_made;
  else if (s == "graph_updated") *value = ctx->graph_updated;
  else if (s == "fp32_rows") *value = ctx->fp32_rows;
  else if (s == "fp32_chunk_rows") *value = ctx->fp32_chunk_rows;
  else if (s == "fp32_streams") *value = ctx->fp32_streams;
  else if (s == "host_threads") *value = ctx->host_threads;
  else if (s == "phase_timing") *value = ctx->phase_timing;
  else return fail(RR_E_ARG, "unknown option " + s);
  return RR_OK;
}

int rr_last_kernel_ms(rr_ctx* ctx, double* ms) {
  if (!ctx || !ms) return fail(RR_E_ARG, "null argument");
  *ms = ctx->last_ms;
  return RR_OK;
}

int rr_last_phase_ms(rr_ctx* ctx, int32_t phase, double* ms) {
  if (!ctx || !ms) return fail(RR_E_ARG, "null argument");
  if (phase < 0 || phase >= RR_PHASES) return fail(RR_E_ARG, "bad phase");
  *ms = ctx->phase_ms[phase];
  return RR_OK;
}

}  // extern "C"
