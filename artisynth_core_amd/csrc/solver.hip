// Solver handle, plan construction and the C ABI (include/kkt_mi355x.h).
//
// Semantics follow the reference's cuDSS bridge (cudssBridge.cc):
//   set_pattern -> analyze -> factor (first: full, later: numeric only)
//   -> solve / iterative_solve, with the state machine UNSET -> ANALYZED ->
//   FACTORED (CuDssSolver.java:57-59), errors as negative status codes and a
//   per-handle last-error string.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <initializer_list>
#include <utility>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "kkt_mi355x.h"
#include "symbolic.h"

using namespace kkt;

namespace {

bool g_timing = false;

// Scratch of kkt_solve_device_noalloc: one 64 MiB region per device, held by
// every device that loads the library until the process ends.  Host code never
// uses the symbol itself (on the host it is only a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per device
// with that device current, and is kept in g_scratch_ptr.  All calls on one
// device share the region, so overlapping calls on one device race.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on device dev (looked up on first use).
hipError_t scratch_on_device(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != dev && (e = hipSetDevice(dev)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Handles may be driven from several host threads (the serving loop does).
// Allocation, free and synchronous copies issued by one thread while another
// thread has a stream capture open are rejected by HIP and can invalidate the
// capture, so captures hold this lock exclusively and every allocation, free
// and synchronous copy holds it shared.
std::shared_mutex g_capture_mu;
typedef std::shared_lock<std::shared_mutex> DevOpLock;

std::string hip_msg(const char* what, hipError_t e) {
  return std::string(what) + ": " + hipGetErrorString(e);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Launch {
  long long off = 0;
  int cnt = 0;
};
// one panel step of a tree level: look-ahead update of the panel's pivot
// columns by the earlier panels of its group, diagonal block, TRSM into W-buffer
// slot `slot`, and (after a group's last panel) the group's trailing update
struct Step {
  Launch look, diag, xrow, trsm, upd;
  int slot = 0, pfirst = 0;
  double upd_flops = 0;
};

// timing classes
enum { TC_ASM = 0, TC_DIAG, TC_TRSM, TC_UPD, TC_SLEVEL, TC_SGEMV, TC_SPMV, TC_VEC, TC_N };
// (TC_SLEVEL: forward solve launches, TC_SGEMV: backward solve launches)

}  // namespace

struct kkt_solver {
  // ---- state ----
  enum { UNSET = 0, PATTERN = 1, ANALYZED = 2, FACTORED = 3 };
  int state = UNSET;
  int n = 0, nnz = 0, mtype = KKT_MT_SYMMETRIC;
  std::vector<int> rowOffs, colIdxs;
  std::vector<int> diagPos;  // position of the diagonal entry of each row, -1 if absent
  int ordering = 1;
  int max_super = KMAX;  // widest supernode of the next analysis
  int prec = KKT_PREC_FP64;   // requested precision of the next factor
  int fprec = KKT_PREC_FP64;  // precision of the current factor
  bool delay_known = false;
  double refine_tol = 1e-10;
  int refine_max = 10;
  int ir_steps = 0;
  int last_refine_steps = 0;
  int fp64_fallbacks = 0;
  bool fallback_active = false;  // the current FP64 factor replaced a stalled FP32 one
  double last_relres = 0.0;
  std::string err;
  bool has_err = false;
  Symbolic S;
  FactorCounters ctr{};

  // ---- device ----
  bool dev_ok = false;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;  // structure/plan allocations (freed on re-analysis)
  // pattern (full symmetric CSR for SpMV) and values
  int *d_rowptr_f = nullptr, *d_col_f = nullptr, *d_map_f = nullptr;
  long long nnzf = 0;
  double *d_vals = nullptr, *d_valsf = nullptr;
  // structure
  DevSym ds{};
  int* d_perm = nullptr;
  int *d_lperm = nullptr;
  signed char* d_piv = nullptr;
  // plan
  int4 *d_asm = nullptr, *d_upd = nullptr;
  int *d_diag = nullptr, *d_aidx = nullptr;
  int2 *d_trsm = nullptr, *d_xrow = nullptr;
  SolveItem *d_fg = nullptr, *d_fr = nullptr, *d_bc = nullptr, *d_ff = nullptr;
  int *d_ilperm = nullptr, *d_gptr = nullptr, *d_gsrc = nullptr;
  long long* d_goff = nullptr;
  int4 *d_zt = nullptr, *d_gf = nullptr;
  std::vector<Launch> xrow_l;  // inverse-row post-pass, one launch per panel index p >= 1
  Launch zt_all, gf_all;       // Z^T and G-form post-passes, one launch each
  long long* d_apos = nullptr;
  // per level: assembly, scatter; forward gather / rows, backward columns
  std::vector<Launch> asm_l, scat_l, fg_l, fr_l, bc_l, ff_l, fs_l;
  std::vector<std::vector<Step>> steps;
  long long front_elems = 0, w_elems = 0, u_elems = 0, v_elems = 0;
  // numeric buffers (allocated lazily per precision)
  double* d_F64 = nullptr;
  float* d_F32 = nullptr;
  // packed solve arrays per precision: forward rows, then (pack_elems + 16
  // later) backward columns, then the forward update slots (sl_elems; zeroed
  // when allocated)
  double* d_PK64 = nullptr;
  float* d_PK32 = nullptr;
  // half-precision copy of the FP32 forward rows and backward columns (same
  // offsets; the FP32 array's slots serve both), read by the iterative solve's
  // preconditioner: half_on = the user setting, have16 = valid for the factor
  _Float16* d_PK16 = nullptr;
  bool half_on = false, have16 = false;
  long long pack_elems = 0, sl_elems = 0;
  int2* d_pack = nullptr;  // packing work items (supernode, block of 4 columns)
  int pack_cnt = 0;
  template <typename T>
  T* pk_fwd(T* PK) const { return PK; }
  template <typename T>
  T* pk_bwd(T* PK) const { return PK + pack_elems + 16; }
  template <typename T>
  T* pk_slot(T* PK) const { return PK + 2 * (pack_elems + 16); }
  size_t pk_elems() const { return (size_t)2 * (pack_elems + 16) + (size_t)sl_elems + 16; }
  size_t pk16_elems() const { return (size_t)2 * (pack_elems + 16); }
  bool use16() const { return have16 && fprec == KKT_PREC_FP32_REFINE; }
  // single-solve vectors, as T: bp, fp, y, xp (n each), w (v_elems + 16)
  size_t svec_elems() const { return (size_t)4 * n + (size_t)v_elems + 16; }
  double* d_W64 = nullptr;
  float* d_W32 = nullptr;
  // solve scratch (svec_elems() entries of the front precision)
  void* d_svec = nullptr;
  size_t svec_bytes = 0;
  // multi-RHS solves: the solve vectors of SOLVE_NRHS right-hand sides (as T);
  // FP64 n-vectors: graph input and output (SOLVE_NRHS each), then residuals,
  // corrections, host-path right-hand sides and solutions (MULTI_BLOCK each);
  // MULTI_BLOCK dot products and their pinned host copy
  static constexpr int MULTI_BLOCK = 64;
  void* d_msvec = nullptr;
  size_t msvec_bytes = 0;
  double* d_mvec = nullptr;
  size_t mvec_bytes = 0;
  double* d_mdot = nullptr;
  double* h_mdot = nullptr;
  double* d_mpart = nullptr;
  // FP64 vectors for refinement / BiCGStab / host copies
  double* d_vec = nullptr;  // 14 vectors of n (0-8 BiCGStab, 10-11 host copies, 12-13 graph I/O)
  double* d_partial = nullptr;
  double* d_part2 = nullptr;  // fused BiCGStab partials: 2 x spmv_blocks(n)
  double* d_scal = nullptr;  // 16 scalars
  double* h_pin = nullptr;   // pinned host copy of the 16 scalars
  // device-side value assembly: scatter plans keyed by the caller's slot list
  struct AddPlan {
    std::vector<int> slots;  // the slot list the plan was built for
    int nuniq = 0;           // distinct valid slots
    int *d_order = nullptr, *d_useg = nullptr, *d_uslot = nullptr;
    unsigned long long used = 0;
  };
  std::vector<AddPlan> add_plans;
  unsigned long long add_clock = 0;
  double* d_contrib = nullptr;  // per-call contributions
  size_t contrib_cap = 0;       // elements
  char* d_stage = nullptr;      // per-call uploaded inputs
  size_t stage_cap = 0;         // bytes
  FactorCounters* d_ctr = nullptr;
  // connected components (assembly trees): per-component max|a| for the pivot
  // perturbation threshold and per-component inertia counters
  int nroots = 0, nachunks = 0;
  long long *d_achunk_a = nullptr, *d_achunk_b = nullptr;
  int* d_acptr = nullptr;
  double *d_amax = nullptr, *d_apart = nullptr;
  int* d_rctr = nullptr;
  std::vector<int> h_rctr;         // after a factor: 4 counters per component
  std::vector<int> root_of_row;    // component of each (original) row
  std::vector<double> multiB;
  // captured triangular-solve graphs, replayed per solve, by solve_q(): 0 FP64,
  // 1 FP32, 2 FP32 fronts with the half-precision copy (the preconditioner)
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t fexec[2] = {nullptr, nullptr};  // captured factorizations
  // captured BiCGStab half-iterations by solve_q(true): 0 head, 1 tail
  hipGraphExec_t bexec[3][2] = {{nullptr, nullptr}, {nullptr, nullptr}, {nullptr, nullptr}};
  // captured multi-RHS triangular solves per precision and count of right-hand sides
  hipGraphExec_t mexec[2][SOLVE_NRHS] = {};
  bool graphs_on = true;
  void drop_graphs() {
    for (int q = 0; q < 3; q++) {
      if (gexec[q]) { (void)hipGraphExecDestroy(gexec[q]); gexec[q] = nullptr; }
      for (int w = 0; w < 2; w++)
        if (bexec[q][w]) { (void)hipGraphExecDestroy(bexec[q][w]); bexec[q][w] = nullptr; }
    }
    for (int q = 0; q < 2; q++) {
      if (fexec[q]) { (void)hipGraphExecDestroy(fexec[q]); fexec[q] = nullptr; }
      for (int w = 0; w < SOLVE_NRHS; w++)
        if (mexec[q][w]) { (void)hipGraphExecDestroy(mexec[q][w]); mexec[q][w] = nullptr; }
    }
  }
  // Captures `fn` (launches on st only) into exec, or returns false.
  template <typename Fn>
  bool capture(hipGraphExec_t& exec, Fn fn) {
    hipGraph_t g = nullptr;
    std::unique_lock<std::shared_mutex> lk(g_capture_mu);
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
    fn();
    if (hipStreamEndCapture(st, &g) != hipSuccess || !g ||
        hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      exec = nullptr;
      (void)hipGetLastError();
      return false;
    }
    (void)hipGraphDestroy(g);
    return true;
  }

  // ---- timing ----
  bool tm_on = false;
  int tm_mask = 0xFF;  // kernel classes that get events
  bool tm_open = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_class;
  size_t ev_used = 0;
  double tm_ms[TC_N] = {0};
  long long tm_cnt[TC_N] = {0};
  double tm_upd_flops = 0;

  kkt_solver() {}
  ~kkt_solver() { release_all(); }

  int fail(int code, const std::string& msg) {
    err = msg;
    has_err = true;
    return code;
  }
  void ok() { has_err = false; }

  void free_list(std::vector<void*>& v) {
    DevOpLock lk(g_capture_mu);
    for (void* p : v)
      if (p) (void)hipFree(p);
    v.clear();
    if (&v == &allocs) alloc_sizes.clear();
  }
  // ---- device memory accounting (bytes) ----
  std::vector<size_t> alloc_sizes;  // parallel to allocs
  size_t plan_bytes() const {
    size_t t = 0;
    for (size_t b : alloc_sizes) t += b;
    return t;
  }
  size_t pattern_bytes() const {
    return (size_t)(n + 1) * 4 + (size_t)std::max<long long>(nnzf, 1) * (4 + 4 + 8) + (size_t)std::max(nnz, 1) * 8 +
           (size_t)14 * n * 8 + (size_t)2 * std::max(spmv_blocks(n), 1) * 8;
  }
  // fronts, D^-1, panel workspace and solve vectors of a factor in precision p
  size_t numeric_bytes(int p) const {
    const size_t esz = (p == KKT_PREC_FP64) ? 8 : 4;
    return ((size_t)front_elems + 2 * (size_t)n + 4) * esz + (size_t)std::max<long long>(w_elems, 1) * esz +
           pk_elems() * esz + svec_elems() * 8 + ((p == KKT_PREC_FP32_REFINE && half_on) ? pk16_elems() * 2 : 0);
  }
  size_t held_bytes() const {
    if (!dev_ok || state < ANALYZED) return 0;
    size_t t = plan_bytes() + pattern_bytes();
    const size_t f64 = ((size_t)front_elems + 2 * (size_t)n + 4 + (size_t)std::max<long long>(w_elems, 1)) * 8;
    const size_t f32 = ((size_t)front_elems + 2 * (size_t)n + 4 + (size_t)std::max<long long>(w_elems, 1)) * 4;
    if (d_F64) t += f64 + pk_elems() * 8;
    if (d_F32) t += f32 + pk_elems() * 4;
    if (d_PK16) t += pk16_elems() * 2;
    t += svec_bytes + msvec_bytes + mvec_bytes;
    return t;
  }
  void release_numeric() {
    drop_graphs();
    DevOpLock lk(g_capture_mu);
    if (d_F64) (void)hipFree(d_F64);
    if (d_F32) (void)hipFree(d_F32);
    if (d_PK64) (void)hipFree(d_PK64);
    if (d_PK32) (void)hipFree(d_PK32);
    if (d_PK16) (void)hipFree(d_PK16);
    d_PK64 = nullptr;
    d_PK32 = nullptr;
    d_PK16 = nullptr;
    have16 = false;
    if (d_W64) (void)hipFree(d_W64);
    if (d_W32) (void)hipFree(d_W32);
    if (d_svec) (void)hipFree(d_svec);
    if (d_msvec) (void)hipFree(d_msvec);
    if (d_mvec) (void)hipFree(d_mvec);
    if (d_mdot) (void)hipFree(d_mdot);
    if (h_mdot) (void)hipHostFree(h_mdot);
    if (d_mpart) (void)hipFree(d_mpart);
    d_mpart = nullptr;
    d_F64 = nullptr; d_F32 = nullptr; d_W64 = nullptr; d_W32 = nullptr; d_svec = nullptr;
    d_msvec = nullptr; d_mvec = nullptr; d_mdot = nullptr; h_mdot = nullptr;
    svec_bytes = msvec_bytes = mvec_bytes = 0;
  }
  void free_plan(AddPlan& p) {
    DevOpLock lk(g_capture_mu);
    if (p.d_order) (void)hipFree(p.d_order);
    if (p.d_useg) (void)hipFree(p.d_useg);
    if (p.d_uslot) (void)hipFree(p.d_uslot);
    p.d_order = p.d_useg = p.d_uslot = nullptr;
  }
  void clear_add_state() {
    for (AddPlan& p : add_plans) free_plan(p);
    add_plans.clear();
    DevOpLock lk(g_capture_mu);
    if (d_contrib) (void)hipFree(d_contrib);
    if (d_stage) (void)hipFree(d_stage);
    d_contrib = nullptr;
    d_stage = nullptr;
    contrib_cap = stage_cap = 0;
  }
  void release_all() {
    if (dev_ok) {
      if (st) (void)hipStreamSynchronize(st);
      clear_add_state();
      free_list(allocs);
      release_numeric();
      DevOpLock lk(g_capture_mu);
      if (d_rowptr_f) (void)hipFree(d_rowptr_f);
      if (d_col_f) (void)hipFree(d_col_f);
      if (d_map_f) (void)hipFree(d_map_f);
      if (d_vals) (void)hipFree(d_vals);
      if (d_valsf) (void)hipFree(d_valsf);
      if (d_vec) (void)hipFree(d_vec);
      if (d_partial) (void)hipFree(d_partial);
      if (d_part2) (void)hipFree(d_part2);
      if (d_scal) (void)hipFree(d_scal);
      if (d_ctr) (void)hipFree(d_ctr);
      if (h_pin) (void)hipHostFree(h_pin);
      h_pin = nullptr;
      for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
      ev_pool.clear();
      for (hipEvent_t e : lev_ev) (void)hipEventDestroy(e);
      lev_ev.clear();
      if (st) (void)hipStreamDestroy(st);
    }
    d_rowptr_f = d_col_f = d_map_f = nullptr;
    d_vals = d_valsf = d_vec = d_partial = d_scal = d_part2 = nullptr;
    d_ctr = nullptr;
    st = nullptr;
    dev_ok = false;
  }

  int dev_id = 0;  // the device current when the handle's stream was created
  int ensure_device() {
    if (dev_ok) return KKT_OK;
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt <= 0)
      return fail(KKT_ERR_DEVICE, "no HIP device available");
    if (hipGetDevice(&dev_id) != hipSuccess) return fail(KKT_ERR_DEVICE, "hipGetDevice failed");
    DevOpLock lk(g_capture_mu);
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      return fail(KKT_ERR_INIT, "hipStreamCreate failed");
    if (hipMalloc(&d_partial, 1024 * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_scal, 16 * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_ctr, sizeof(FactorCounters)) != hipSuccess ||
        hipHostMalloc(&h_pin, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return fail(KKT_ERR_ALLOC, "hipMalloc failed (scalars)");
    dev_ok = true;
    return KKT_OK;
  }

  template <typename T>
  int upload(T*& dptr, const std::vector<T>& h) {
    size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
    DevOpLock lk(g_capture_mu);
    hipError_t e = hipMalloc(&dptr, bytes);
    if (e != hipSuccess) return fail(KKT_ERR_ALLOC, hip_msg("hipMalloc failed (plan)", e));
    allocs.push_back(dptr);
    alloc_sizes.push_back(bytes);
    if (!h.empty() &&
        (e = hipMemcpy(dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice)) != hipSuccess)
      return fail(KKT_ERR_COPY, hip_msg("hipMemcpy failed (plan)", e));
    return KKT_OK;
  }

  // ---- timing helpers ----
  void tbeg(int cls) {
    tm_open = false;
    if (!tm_on || !((tm_mask >> cls) & 1)) return;
    tm_open = true;
    if (ev_used + 2 > ev_pool.size()) {
      DevOpLock lk(g_capture_mu);
      for (int q = 0; q < 256; q++) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        ev_pool.push_back(e);
      }
    }
    (void)hipEventRecord(ev_pool[ev_used], st);
    ev_class.push_back(cls);
    ev_used++;
  }
  void tend() {
    if (!tm_open) return;
    tm_open = false;
    (void)hipEventRecord(ev_pool[ev_used], st);
    ev_used++;
  }
  void tflush() {  // after a stream sync
    if (!tm_on) return;
    for (size_t q = 0; q + 1 < ev_used; q += 2) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, ev_pool[q], ev_pool[q + 1]);
      int c = ev_class[q / 2];
      tm_ms[c] += ms;
      tm_cnt[c]++;
    }
    ev_used = 0;
    ev_class.clear();
  }

  // ------------------------------------------------------------------
  int set_pattern(int n_, int nnz_, const int* ro, const int* ci, int mt) {
    if (mt != KKT_MT_SYMMETRIC && mt != KKT_MT_SPD) {
      if (mt == KKT_MT_GENERAL)
        return fail(KKT_ERR_UNKNOWN_MTYPE, "general (unsymmetric) matrices are not supported; "
                                           "use a symmetric type");
      return fail(KKT_ERR_UNKNOWN_MTYPE, "unknown matrix type");
    }
    std::string e;
    if (validate_pattern(n_, nnz_, ro, ci, e) != 0) return fail(KKT_ERR_PATTERN, e);
    if (dev_ok) {
      (void)hipStreamSynchronize(st);
      free_list(allocs);
      release_numeric();
      DevOpLock lk(g_capture_mu);
      if (d_rowptr_f) (void)hipFree(d_rowptr_f);
      if (d_col_f) (void)hipFree(d_col_f);
      if (d_map_f) (void)hipFree(d_map_f);
      if (d_vals) (void)hipFree(d_vals);
      if (d_valsf) (void)hipFree(d_valsf);
      if (d_vec) (void)hipFree(d_vec);
      if (d_part2) (void)hipFree(d_part2);
      d_rowptr_f = d_col_f = d_map_f = nullptr;
      d_vals = d_valsf = d_vec = d_part2 = nullptr;
    }
    n = n_;
    nnz = nnz_;
    mtype = mt;
    rowOffs.assign(ro, ro + n + 1);
    colIdxs.assign(ci, ci + nnz);
    diagPos.assign(n, -1);
    for (int i = 0; i < n; i++)
      for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++)
        if (colIdxs[p] == i) diagPos[i] = p;
    state = PATTERN;
    delay_known = false;
    return KKT_OK;
  }

  int analyze_with_mask(const unsigned char* delay) {
    SymOptions opt;
    opt.ordering = ordering;
    opt.max_super = max_super;
    std::string e;
    double t0 = now_ms();
    int rc = symbolic_analyze(n, rowOffs.data(), colIdxs.data(), delay, opt, S, e);
    if (rc != 0) return fail(KKT_ERR_ANALYSIS, "symbolic analysis failed: " + e);
    double t1 = now_ms();
    rc = ensure_device();
    if (rc) return rc;
    rc = build_plan();
    if (rc) return rc;
    if (g_timing)
      fprintf(stderr, "[kkt-timing] analyze n=%d nnz=%d: symbolic %.2f ms, plan+upload %.2f ms, "
              "nsuper=%d levels=%d nnzL=%lld flops=%.3g\n",
              n, nnz, t1 - t0, now_ms() - t1, S.nsuper, S.nlevels, (long long)S.nnzL, S.flops);
    state = ANALYZED;
    return KKT_OK;
  }

  int analyze(const double* hvals) {
    if (state == UNSET) return fail(KKT_ERR_STATE, "analyze called before setPattern");
    if (hvals) {
      std::vector<unsigned char> mask;
      make_mask(hvals, mask);
      delay_known = true;
      return analyze_with_mask(mask.data());
    }
    delay_known = false;
    return analyze_with_mask(nullptr);
  }

  // zero (or absent) diagonal -> constraint-like row ordered after its neighbours
  bool make_mask(const double* hvals, std::vector<unsigned char>& mask) {
    mask.assign(n, 0);
    bool any = false;
    for (int i = 0; i < n; i++) {
      int p = diagPos[i];
      if (p < 0 || hvals[p] == 0.0) { mask[i] = 1; any = true; }
    }
    return any;
  }

  // ---- solve work items ----
  // Per level, the forward-rows and backward-columns kernels give each row or
  // column 2^lg lanes, the smallest power of two with epl * 2^lg >= its length
  // (rows: up to 64 lanes; columns: up to 256, several waves).  epl (elements
  // per lane) is the shape the autotuner picks per level; the gather items do
  // not depend on it.
  std::vector<int> sv_k, sv_m, sv_first;
  std::vector<long long> sv_foff, sv_uoff, sv_roff, sv_goff, sv_voff, sv_poff, sv_soff, sv_sdst;
  std::vector<int> sv_nch;
  std::vector<std::vector<int>> sv_lv;
  std::vector<int> fr_epl, bc_epl;
  std::vector<int> fr_fuse;  // per level: 1 = one forward launch (kind 2 items), 0 = gather + rows
  std::vector<int> fr_rg;    // per level: row groups per kind-2 item (sharing one gather)
  // Row groups per forward row block, per level: the most (up to eight) that
  // still leave about 1024 row blocks in the level's launch, so every CU keeps
  // several workgroups while the blocks of a front gather its right-hand side
  // fewer times.
  std::vector<int> default_row_groups(int epl) const {
    const int nl = (int)sv_lv.size();
    std::vector<int> rg(nl, 1);
    for (int L = 0; L < nl; L++) {
      long long blocks = 0;
      for (int s : sv_lv[L]) {
        const int k = sv_k[s], m = sv_m[s];
        if (k <= FWD_SMALL_K && m <= FWD_SMALL_M) continue;
        int lg = 0;
        while (lg < 6 && epl * (1 << lg) < k) lg++;
        const int rf = 4 * (64 >> lg);
        blocks += (m + rf - 1) / rf;
      }
      while (rg[L] < 8 && blocks / (2 * rg[L]) >= 1024) rg[L] *= 2;
    }
    return rg;
  }
  int replace_upload(SolveItem*& dptr, const std::vector<SolveItem>& v) {
    // a pointer no longer in allocs was already freed with them (new analysis)
    if (dptr) {
      DevOpLock lk(g_capture_mu);
      auto it = std::find(allocs.begin(), allocs.end(), (void*)dptr);
      if (it != allocs.end()) {
        alloc_sizes.erase(alloc_sizes.begin() + (it - allocs.begin()));
        allocs.erase(it);
        (void)hipFree(dptr);
      }
      dptr = nullptr;
    }
    return upload(dptr, v);
  }
  int build_solve_items() {
    const int nl = (int)sv_lv.size();
    // fg / fr: the two forward launches per level of the multi-RHS solve (gather,
    // then rows).  ff: the forward launches of the single solve, per level either
    // one (small fronts whole, row blocks of large fronts with their own gather,
    // kind 2) or two (small fronts and gather blocks, kinds 0 and 1, then the
    // row blocks, kind 3: the fs range)
    std::vector<SolveItem> fgv, frv, bcv, ffv, fsv;
    if ((int)fr_fuse.size() != nl) fr_fuse.assign(nl, 1);
    if ((int)fr_rg.size() != nl) fr_rg.assign(nl, 1);
    auto item = [&](int s, int blk, int lg, int kind) {
      SolveItem it;
      it.s = s; it.blk = blk; it.lg = lg; it.kind = kind;
      it.k = sv_k[s]; it.m = sv_m[s]; it.first = sv_first[s]; it.pad = 0;
      it.foff = sv_foff[s]; it.uoff = sv_uoff[s]; it.roff = sv_roff[s]; it.goff = sv_goff[s];
      it.voff = sv_voff[s]; it.poff = sv_poff[s];
      it.soff = sv_soff[s]; it.sdst = sv_sdst[s]; it.nch = sv_nch[s];
      it.pad1 = it.pad2 = it.pad3 = 0;
      return it;
    };
    auto seg_lg = [](int len, int epl, int cap) {
      int lg = 0;
      while (lg < cap && epl * (1 << lg) < len) lg++;
      return lg;
    };
    fg_l.assign(nl, Launch());
    fr_l.assign(nl, Launch());
    bc_l.assign(nl, Launch());
    ff_l.assign(nl, Launch());
    fs_l.assign(nl, Launch());
    for (int L = 0; L < nl; L++) {
      fsv.clear();
      fg_l[L].off = (long long)fgv.size();
      fr_l[L].off = (long long)frv.size();
      bc_l[L].off = (long long)bcv.size();
      ff_l[L].off = (long long)ffv.size();
      for (int s : sv_lv[L]) {
        const int k = sv_k[s], m = sv_m[s];
        if (k > KMAX) return fail(KKT_ERR_ANALYSIS, "supernode wider than the solve kernels take");
        if (k <= FWD_SMALL_K && m <= FWD_SMALL_M) {
          // small front: gather and rows in one workgroup (the first forward launch)
          fgv.push_back(item(s, 0, 0, 0));
          ffv.push_back(item(s, 0, 0, 0));
        } else {
          for (int gb = 0; gb < (m + 255) / 256; gb++) {
            fgv.push_back(item(s, gb, 0, 1));
            if (!fr_fuse[L]) ffv.push_back(item(s, gb, 0, 1));
          }
          const int lf = seg_lg(k, fr_epl[L], 6), rf = 4 * (64 >> lf);
          for (int rb = 0; rb < (m + rf - 1) / rf; rb++) {
            frv.push_back(item(s, rb, lf, 0));
            if (!fr_fuse[L]) fsv.push_back(item(s, rb, lf, 3));
          }
          if (fr_fuse[L]) {
            // one-launch form: row blocks of nrg groups of rf rows share a gather
            int nrg = std::max(1, fr_rg[L]);
            while (nrg > 1 && nrg * rf > FWD_GROUP_ROWS) nrg >>= 1;
            for (int rb = 0; rb < (m + nrg * rf - 1) / (nrg * rf); rb++) {
              SolveItem it = item(s, rb, lf, 2);
              it.pad = nrg;
              ffv.push_back(it);
            }
          }
        }
        const int lb = seg_lg(m, bc_epl[L], 8), rbw = 256 >> lb;
        for (int cb = 0; cb < (k + rbw - 1) / rbw; cb++) bcv.push_back(item(s, cb, lb, 0));
      }
      fg_l[L].cnt = (int)(fgv.size() - fg_l[L].off);
      fr_l[L].cnt = (int)(frv.size() - fr_l[L].off);
      bc_l[L].cnt = (int)(bcv.size() - bc_l[L].off);
      ff_l[L].cnt = (int)(ffv.size() - ff_l[L].off);
      fs_l[L].off = (long long)ffv.size();
      fs_l[L].cnt = (int)fsv.size();
      ffv.insert(ffv.end(), fsv.begin(), fsv.end());
    }
    int rc;
    if ((rc = replace_upload(d_fg, fgv)) || (rc = replace_upload(d_fr, frv)) || (rc = replace_upload(d_bc, bcv)) ||
        (rc = replace_upload(d_ff, ffv)))
      return rc;
    return KKT_OK;
  }

  // ------------------------------------------------------------------
  int build_plan() {
    (void)hipStreamSynchronize(st);
    clear_add_state();
    free_list(allocs);
    release_numeric();
    {
      DevOpLock lk(g_capture_mu);
      for (void* q : {(void*)d_rowptr_f, (void*)d_col_f, (void*)d_map_f, (void*)d_vals, (void*)d_valsf,
                      (void*)d_vec, (void*)d_part2})
        if (q) (void)hipFree(q);
    }
    d_rowptr_f = d_col_f = d_map_f = nullptr;
    d_vals = d_valsf = d_vec = d_part2 = nullptr;
    const int ns = S.nsuper;
    std::vector<int> first(ns), kk(ns), mm(ns);
    std::vector<long long> foff(ns + 1), roff(ns + 1), woff(ns + 1), uoff(ns + 1), voff(ns + 1);
    foff[0] = roff[0] = woff[0] = uoff[0] = voff[0] = 0;
    for (int s = 0; s < ns; s++) {
      first[s] = S.sup_ptr[s];
      kk[s] = S.sup_ptr[s + 1] - S.sup_ptr[s];
      long long r = S.rows_ptr[s + 1] - S.rows_ptr[s];
      mm[s] = kk[s] + (int)r;
      foff[s + 1] = foff[s] + (long long)mm[s] * mm[s];
      roff[s + 1] = roff[s] + r;
      woff[s + 1] = woff[s] + (long long)UPD_GROUP * mm[s] * NB;
      uoff[s + 1] = uoff[s] + r;
      voff[s + 1] = voff[s] + mm[s];
    }
    for (int s = 0; s < ns; s++)
      if (kk[s] > KMAX) return fail(KKT_ERR_ANALYSIS, "internal: supernode wider than the solve kernels take");
    front_elems = foff[ns];
    w_elems = woff[ns];
    u_elems = uoff[ns];
    v_elems = voff[ns];
    if (v_elems >= (1LL << 31) - 16) return fail(KKT_ERR_ANALYSIS, "internal: solve vector too large");
    // children (ascending)
    std::vector<int> chp(ns + 1, 0), chl;
    for (int s = 0; s < ns; s++)
      if (S.sparent[s] >= 0) chp[S.sparent[s] + 1]++;
    for (int s = 0; s < ns; s++) chp[s + 1] += chp[s];
    chl.assign(std::max(chp[ns], 1), 0);
    {
      std::vector<int> pos(chp.begin(), chp.end() - 1);
      for (int s = 0; s < ns; s++)
        if (S.sparent[s] >= 0) chl[pos[S.sparent[s]]++] = s;
    }
    // relative positions of each child's rows in its parent front
    std::vector<int> rel(std::max<long long>(roff[ns], 1), 0);
    for (int c = 0; c < ns; c++) {
      int p = S.sparent[c];
      if (p < 0) continue;
      int fp = first[p], kp = kk[p];
      const int* prow = S.rows.data() + S.rows_ptr[p];
      int rp = (int)(S.rows_ptr[p + 1] - S.rows_ptr[p]);
      int q = 0;
      for (long long i = S.rows_ptr[c]; i < S.rows_ptr[c + 1]; i++) {
        int g = S.rows[i];
        int pos;
        if (g < fp + kp) pos = g - fp;
        else {
          while (q < rp && prow[q] < g) q++;
          if (q >= rp || prow[q] != g) return fail(KKT_ERR_ANALYSIS, "internal: child row not in parent");
          pos = kp + q;
        }
        rel[i] = pos;
      }
    }
    // forward gather map: front position -> the children's update entries, in
    // ascending child order (the summation order of the forward sweep)
    std::vector<long long> goff(ns + 1, 0);
    for (int s = 0; s < ns; s++) goff[s + 1] = goff[s] + mm[s];
    std::vector<int> gptr(goff[ns] + 1, 0), gsrc;
    for (int s = 0; s < ns; s++)
      for (int q = chp[s]; q < chp[s + 1]; q++) {
        const int c = chl[q];
        for (long long i = roff[c]; i < roff[c + 1]; i++) gptr[goff[s] + rel[i] + 1]++;
      }
    for (long long g = 0; g < goff[ns]; g++) {
      if ((long long)gptr[g] + gptr[g + 1] > INT32_MAX) return fail(KKT_ERR_ANALYSIS, "gather map too large");
      gptr[g + 1] += gptr[g];
    }
    gsrc.assign(std::max(gptr[goff[ns]], 1), 0);
    {
      std::vector<int> fill(gptr.begin(), gptr.end() - 1);
      for (int s = 0; s < ns; s++)
        for (int q = chp[s]; q < chp[s + 1]; q++) {
          const int c = chl[q];
          for (long long i = roff[c]; i < roff[c + 1]; i++)
            gsrc[fill[goff[s] + rel[i]]++] = (int)(uoff[c] + (i - roff[c]));
        }
    }
    // forward update slots: an nch x m block per supernode (children in the
    // ascending order above), and where each child's contribution rows go
    std::vector<long long> soff(ns + 1, 0), sdst(std::max(ns, 1), 0);
    std::vector<int> nchv(std::max(ns, 1), 0);
    for (int s = 0; s < ns; s++) {
      nchv[s] = chp[s + 1] - chp[s];
      soff[s + 1] = soff[s] + (long long)nchv[s] * mm[s];
      for (int q = chp[s]; q < chp[s + 1]; q++) sdst[chl[q]] = soff[s] + (long long)(q - chp[s]) * mm[s];
    }
    sl_elems = soff[ns];
    // levels
    int nl = S.nlevels;
    std::vector<std::vector<int>> lv(nl);
    for (int s = 0; s < ns; s++) lv[S.level[s]].push_back(s);
    // A entries -> fronts, bucketed by level
    std::vector<std::vector<int>> a_idx(nl);
    std::vector<std::vector<long long>> a_pos(nl);
    for (int i = 0; i < n; i++)
      for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
        int j = colIdxs[p];
        if (j < i) continue;
        int a = S.iperm[i], b = S.iperm[j];
        int c = std::min(a, b), r = std::max(a, b);
        int s = S.col2sup[c];
        int fs = first[s], k = kk[s], m = mm[s];
        int rp;
        if (r < fs + k) rp = r - fs;
        else {
          const int* rw = S.rows.data() + S.rows_ptr[s];
          int cnt = (int)(S.rows_ptr[s + 1] - S.rows_ptr[s]);
          int q = (int)(std::lower_bound(rw, rw + cnt, r) - rw);
          if (q >= cnt || rw[q] != r) return fail(KKT_ERR_ANALYSIS, "internal: entry outside structure");
          rp = k + q;
        }
        int L = S.level[s];
        a_idx[L].push_back(p);
        a_pos[L].push_back(foff[s] + (long long)(c - fs) * m + rp);
      }
    // work items
    std::vector<int4> asmv, updv;
    std::vector<int> diagv, aidx;
    std::vector<long long> apos;
    std::vector<int2> trsmv, xrowv;
    asm_l.assign(nl, Launch());
    scat_l.assign(nl, Launch());
    steps.assign(nl, std::vector<Step>());
    for (int L = 0; L < nl; L++) {
      asm_l[L].off = (long long)asmv.size();
      for (int s : lv[L]) {
        int t = (mm[s] + TILE - 1) / TILE;
        for (int I = 0; I < t; I++)
          for (int J = 0; J <= I; J++) asmv.push_back(make_int4(s, I, J, 0));
      }
      asm_l[L].cnt = (int)(asmv.size() - asm_l[L].off);
      scat_l[L].off = (long long)aidx.size();
      aidx.insert(aidx.end(), a_idx[L].begin(), a_idx[L].end());
      apos.insert(apos.end(), a_pos[L].begin(), a_pos[L].end());
      scat_l[L].cnt = (int)a_idx[L].size();
      int P = 0;
      for (int s : lv[L]) P = std::max(P, (kk[s] + NB - 1) / NB);
      steps[L].resize(P);
      // Panels go in groups of UPD_GROUP: the trailing Schur update of a group is
      // one launch with K = the group's columns; inside a group, each panel's
      // pivot columns first receive the updates of the group's earlier panels
      // (look-ahead), so the diagonal block and TRSM see the same values as with
      // an update after every panel.
      for (int p = 0; p < P; p++) {
        Step& sp = steps[L][p];
        const int pf = (p / UPD_GROUP) * UPD_GROUP;
        sp.slot = p - pf;
        sp.pfirst = pf;
        sp.look.off = (long long)updv.size();
        if (p > pf)
          for (int s : lv[L]) {
            const int j0 = p * NB;
            if (kk[s] <= j0) continue;
            const int w = std::min(NB, kk[s] - j0), t = (mm[s] - j0 + TILE - 1) / TILE;
            for (int I = 0; I < t; I++) updv.push_back(make_int4(s, I, -1, (p - pf) | (j0 << 3)));
            sp.upd_flops += 2.0 * (double)((p - pf) * NB) * (double)(mm[s] - j0) * (double)w;
          }
        sp.look.cnt = (int)(updv.size() - sp.look.off);
        sp.diag.off = (long long)diagv.size();
        sp.xrow.off = (long long)xrowv.size();
        sp.trsm.off = (long long)trsmv.size();
        for (int s : lv[L]) {
          int j0 = p * NB;
          if (kk[s] <= j0) continue;
          int w = std::min(NB, kk[s] - j0), r0 = j0 + w, mr = mm[s] - r0;
          diagv.push_back(s);
          for (int rb = 0; rb < (mr + 63) / 64; rb++) trsmv.push_back(make_int2(s, rb));
        }
        sp.diag.cnt = (int)(diagv.size() - sp.diag.off);
        sp.xrow.cnt = (int)(xrowv.size() - sp.xrow.off);
        sp.trsm.cnt = (int)(trsmv.size() - sp.trsm.off);
        sp.upd.off = (long long)updv.size();
        if (p % UPD_GROUP == UPD_GROUP - 1 || p == P - 1)
          for (int s : lv[L]) {
            const int ps = (kk[s] + NB - 1) / NB;
            if (ps <= pf) continue;
            const int np = std::min(ps, pf + UPD_GROUP) - pf;
            const int rs = std::min(kk[s], (pf + np) * NB), mr = mm[s] - rs;
            const int t = (mr + TILE - 1) / TILE;
            for (int I = 0; I < t; I++)
              for (int J = 0; J <= I; J++) updv.push_back(make_int4(s, I, J, np | (rs << 3)));
            sp.upd_flops += (double)(rs - pf * NB) * (double)mr * (double)(mr + 1);
          }
        sp.upd.cnt = (int)(updv.size() - sp.upd.off);
      }
    }
    // Z^T and G-form post-passes over all supernodes
    std::vector<int4> ztv, gfv;
    for (int s = 0; s < ns; s++) {
      const int nc = (kk[s] + 63) / 64, nr = (mm[s] - kk[s] + 63) / 64;
      for (int cb = 0; cb < nc; cb++) {
        for (int jb = 0; jb < nr; jb++) {
          ztv.push_back(make_int4(s, cb, jb, 0));
          gfv.push_back(make_int4(s, cb, jb, 1));
        }
        for (int ib = cb; ib < nc; ib++) gfv.push_back(make_int4(s, ib, cb, 0));
      }
    }
    zt_all.off = gf_all.off = 0;
    zt_all.cnt = (int)ztv.size();
    gf_all.cnt = (int)gfv.size();
    // inverse rows X(P, :) of every pivot block, after the whole factorization:
    // launch p covers block row p of all supernodes at once
    {
      int maxp = 0;
      for (int s = 0; s < ns; s++) maxp = std::max(maxp, (kk[s] + NB - 1) / NB);
      xrow_l.assign(std::max(maxp, 1), Launch());
      for (int p = 1; p < maxp; p++) {
        xrow_l[p].off = (long long)xrowv.size();
        for (int s = 0; s < ns; s++)
          if (kk[s] > p * NB)
            for (int cb = 0; cb < p; cb++) xrowv.push_back(make_int2(s, cb));
        xrow_l[p].cnt = (int)(xrowv.size() - xrow_l[p].off);
      }
    }
    // full symmetric CSR for SpMV (row-sorted), map to upper value index
    std::vector<int> rp_f(n + 1, 0);
    for (int i = 0; i < n; i++)
      for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
        int j = colIdxs[p];
        if (j < i) continue;
        rp_f[i + 1]++;
        if (j != i) rp_f[j + 1]++;
      }
    for (int i = 0; i < n; i++) rp_f[i + 1] += rp_f[i];
    nnzf = rp_f[n];
    std::vector<int> col_f(std::max<long long>(nnzf, 1)), map_f(std::max<long long>(nnzf, 1));
    {
      std::vector<int> pos(rp_f.begin(), rp_f.end() - 1);
      for (int i = 0; i < n; i++)
        for (int p = rowOffs[i]; p < rowOffs[i + 1]; p++) {
          int j = colIdxs[p];
          if (j < i) continue;
          col_f[pos[i]] = j; map_f[pos[i]++] = p;
          if (j != i) { col_f[pos[j]] = i; map_f[pos[j]++] = p; }
        }
      std::vector<std::pair<int, int>> tmp;
      for (int i = 0; i < n; i++) {
        tmp.clear();
        for (int q = rp_f[i]; q < rp_f[i + 1]; q++) tmp.push_back({col_f[q], map_f[q]});
        std::sort(tmp.begin(), tmp.end());
        for (int q = rp_f[i]; q < rp_f[i + 1]; q++) { col_f[q] = tmp[q - rp_f[i]].first; map_f[q] = tmp[q - rp_f[i]].second; }
      }
    }
    // uploads (rows padded: the solve kernels' chunk loads may run 3 entries past
    // a supernode's rows)
    std::vector<int> rows_pad(S.rows);
    rows_pad.resize(S.rows.size() + 4, 0);
    // backward vector: supernode s owns entries [voff_s, voff_s + m_s): its pivots,
    // then its below-diagonal rows.  Push map: for each solve position g, the
    // entries of the rows naming it (ascending); wpos: the entry of each pivot
    std::vector<int> pptr(n + 1, 0), ppos(std::max<size_t>(S.rows.size(), 1), 0), wpos(std::max(n, 1), 0);
    for (size_t e = 0; e < S.rows.size(); e++) pptr[S.rows[e] + 1]++;
    for (int g = 0; g < n; g++) pptr[g + 1] += pptr[g];
    {
      std::vector<int> fill(pptr.begin(), pptr.end() - 1);
      for (int s = 0; s < ns; s++) {
        for (int t = 0; t < kk[s]; t++) wpos[first[s] + t] = (int)(voff[s] + t);
        for (long long j = 0; j < roff[s + 1] - roff[s]; j++)
          ppos[fill[S.rows[roff[s] + j]]++] = (int)(voff[s] + kk[s] + j);
      }
    }
    // connected components: the assembly trees, numbered in ascending order of
    // their root supernode; each CSR row belongs to the tree of its pivot column.
    // Value ranges of consecutive rows of one component are cut into chunks of at
    // most 8192 entries for the per-component max|a| pass.
    std::vector<int> sroot(ns, -1);
    nroots = 0;
    {
      std::vector<int> rid(ns, -1);
      for (int s = 0; s < ns; s++)
        if (S.sparent[s] < 0) rid[s] = nroots++;
      for (int s = ns - 1; s >= 0; s--) sroot[s] = S.sparent[s] < 0 ? rid[s] : sroot[S.sparent[s]];
    }
    root_of_row.assign(n, 0);
    for (int i = 0; i < n; i++) root_of_row[i] = sroot[S.col2sup[S.iperm[i]]];
    std::vector<std::vector<std::pair<long long, long long>>> rranges(std::max(nroots, 1));
    for (int i = 0; i < n;) {
      int j = i;
      while (j + 1 < n && root_of_row[j + 1] == root_of_row[i]) j++;
      rranges[root_of_row[i]].push_back({(long long)rowOffs[i], (long long)rowOffs[j + 1]});
      i = j + 1;
    }
    std::vector<long long> ca, cb;
    std::vector<int> cptr(std::max(nroots, 1) + 1, 0);
    for (int r = 0; r < nroots; r++) {
      for (auto& q : rranges[r])
        for (long long a = q.first; a < q.second; a += 8192) {
          ca.push_back(a);
          cb.push_back(std::min(q.second, a + 8192));
        }
      cptr[r + 1] = (int)ca.size();
    }
    nachunks = (int)ca.size();
    // packed solve arrays: supernodes in level-major order, k(k-1)/2 + k r
    // entries each; packing items (supernode, block of 4 front columns)
    std::vector<long long> poffv(std::max(ns, 1), 0);
    std::vector<int2> packv;
    pack_elems = 0;
    for (int L = 0; L < nl; L++)
      for (int s : lv[L]) {
        poffv[s] = pack_elems;
        const long long k = kk[s], r = mm[s] - kk[s];
        pack_elems += k * (k - 1) / 2 + k * r;
        for (int cb = 0; cb < (mm[s] + 3) / 4; cb++) packv.push_back(make_int2(s, cb));
      }
    pack_cnt = (int)packv.size();
    int rc = 0;
    long long* d_poffp = nullptr;
    if ((rc = upload(d_poffp, poffv)) || (rc = upload(d_pack, packv))) return rc;
    int *d_first, *d_k, *d_m, *d_chp, *d_chl, *d_rel, *d_rows, *d_pptr, *d_ppos, *d_wpos, *d_sroot;
    std::vector<double> zd(std::max(std::max(nachunks, nroots), 1), 0.0);
    std::vector<int> zi(4 * std::max(nroots, 1), 0);
    if ((rc = upload(d_achunk_a, ca)) || (rc = upload(d_achunk_b, cb)) || (rc = upload(d_acptr, cptr)) ||
        (rc = upload(d_sroot, sroot)) || (rc = upload(d_amax, zd)) || (rc = upload(d_apart, zd)) ||
        (rc = upload(d_rctr, zi)))
      return rc;
    long long *d_foff, *d_roff, *d_woff, *d_uoff;
    if ((rc = upload(d_first, first)) || (rc = upload(d_k, kk)) || (rc = upload(d_m, mm)) ||
        (rc = upload(d_foff, foff)) || (rc = upload(d_roff, roff)) || (rc = upload(d_woff, woff)) ||
        (rc = upload(d_uoff, uoff)) || (rc = upload(d_chp, chp)) || (rc = upload(d_chl, chl)) ||
        (rc = upload(d_rel, rel)) || (rc = upload(d_rows, rows_pad)) || (rc = upload(d_perm, S.perm)) ||
        (rc = upload(d_asm, asmv)) || (rc = upload(d_upd, updv)) || (rc = upload(d_diag, diagv)) ||
        (rc = upload(d_aidx, aidx)) || (rc = upload(d_apos, apos)) || (rc = upload(d_trsm, trsmv)) ||
        (rc = upload(d_xrow, xrowv)) || (rc = upload(d_zt, ztv)) || (rc = upload(d_gf, gfv)) ||
        (rc = upload(d_pptr, pptr)) || (rc = upload(d_ppos, ppos)) || (rc = upload(d_wpos, wpos)))
      return rc;
    // solve work items, default shapes (sixteen elements per lane, one forward
    // launch per level)
    sv_k = kk; sv_m = mm; sv_first = first; sv_foff = foff; sv_uoff = uoff; sv_roff = roff;
    sv_goff = goff; sv_voff = voff; sv_lv = lv; sv_poff = poffv;
    sv_soff.assign(soff.begin(), soff.end() - 1); sv_sdst = sdst; sv_nch = nchv;
    fr_epl.assign(nl, 16);
    bc_epl.assign(nl, 16);
    fr_fuse.assign(nl, 1);
    fr_rg = default_row_groups(16);
    if ((rc = build_solve_items())) return rc;
    std::vector<int> lp0(n);
    for (int s = 0; s < ns; s++)
      for (int t = 0; t < kk[s]; t++) lp0[first[s] + t] = t;
    std::vector<signed char> pv0(n, 1);
    if ((rc = upload(d_lperm, lp0)) || (rc = upload(d_ilperm, lp0)) || (rc = upload(d_piv, pv0)) ||
        (rc = upload(d_goff, goff)) || (rc = upload(d_gptr, gptr)) || (rc = upload(d_gsrc, gsrc)))
      return rc;
    ds.sn_first = d_first; ds.sn_k = d_k; ds.sn_m = d_m;
    ds.sn_foff = d_foff; ds.sn_roff = d_roff; ds.sn_woff = d_woff; ds.sn_uoff = d_uoff;
    ds.ch_ptr = d_chp; ds.ch_list = d_chl; ds.rel = d_rel; ds.rows = d_rows;
    ds.push_ptr = d_pptr; ds.push_pos = d_ppos; ds.wpos = d_wpos;
    ds.lperm = d_lperm; ds.ilperm = d_ilperm; ds.piv = d_piv;
    ds.sn_goff = d_goff; ds.gptr = d_gptr; ds.gsrc = d_gsrc;
    ds.sn_root = d_sroot;
    ds.sn_poff = d_poffp;
    // host copies of what was uploaded, for the structure-parity check
    h_first = first; h_k = kk; h_m = mm; h_foff = foff; h_roff = roff; h_chp = chp; h_chl = chl;
    h_rel = rel;
    // pattern + values
    DevOpLock lk(g_capture_mu);
    if (hipMalloc(&d_rowptr_f, (n + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_col_f, col_f.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_map_f, map_f.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_vals, std::max(nnz, 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_valsf, col_f.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_vec, (size_t)14 * n * sizeof(double)) != hipSuccess ||
        hipMalloc(&d_part2, (size_t)2 * std::max(spmv_blocks(n), 1) * sizeof(double)) != hipSuccess)
      return fail(KKT_ERR_ALLOC, "hipMalloc failed (pattern)");
    if (hipMemcpy(d_rowptr_f, rp_f.data(), (n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_col_f, col_f.data(), col_f.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_map_f, map_f.data(), map_f.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      return fail(KKT_ERR_COPY, "hipMemcpy failed (pattern)");
    return KKT_OK;
  }
  std::vector<int> h_first, h_k, h_m, h_chp, h_chl, h_rel;
  std::vector<long long> h_foff, h_roff;

  int ensure_numeric(int p) {
    size_t esz = (p == KKT_PREC_FP64) ? sizeof(double) : sizeof(float);
    DevOpLock lk(g_capture_mu);
    if (p == KKT_PREC_FP64 && !d_F64) {
      // fronts, then the two D^-1 vectors (E0, E1) of n entries each, then 4 spare
      // entries (the solve kernels' 16-byte chunk loads may run 3 past a column)
      if (hipMalloc(&d_F64, (front_elems + 2LL * n + 4) * sizeof(double)) != hipSuccess ||
          hipMalloc(&d_W64, std::max<long long>(w_elems, 1) * sizeof(double)) != hipSuccess ||
          hipMalloc(&d_PK64, pk_elems() * sizeof(double)) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (FP64 fronts)");
      if (hipMemsetAsync(d_PK64, 0, pk_elems() * sizeof(double), st) != hipSuccess)
        return fail(KKT_ERR_DEVICE, "clearing the update slots failed");
    }
    if (p == KKT_PREC_FP32_REFINE && !d_F32) {
      if (hipMalloc(&d_F32, (front_elems + 2LL * n + 4) * sizeof(float)) != hipSuccess ||
          hipMalloc(&d_W32, std::max<long long>(w_elems, 1) * sizeof(float)) != hipSuccess ||
          hipMalloc(&d_PK32, pk_elems() * sizeof(float)) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (FP32 fronts)");
      if (hipMemsetAsync(d_PK32, 0, pk_elems() * sizeof(float), st) != hipSuccess)
        return fail(KKT_ERR_DEVICE, "clearing the update slots failed");
    }
    if (p == KKT_PREC_FP32_REFINE && half_on && !d_PK16) {
      drop_graphs();  // the factor graph gains the conversion
      if (hipMalloc(&d_PK16, pk16_elems() * sizeof(_Float16)) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (half-precision solve copy)");
    }
    size_t need = svec_elems() * std::max(esz, sizeof(double));
    if (svec_bytes < need) {
      drop_graphs();
      if (d_svec) (void)hipFree(d_svec);
      if (hipMalloc(&d_svec, need) != hipSuccess) return fail(KKT_ERR_ALLOC, "hipMalloc failed (solve)");
      svec_bytes = need;
    }
    return KKT_OK;
  }

  // ------------------------------------------------------------------
  template <typename T>
  void run_factor(T* F, T* W, T* PK, double pert) {
    tbeg(TC_VEC);
    launch_absmax_roots(nachunks, d_achunk_a, d_achunk_b, d_vals, d_apart, nroots, d_acptr, d_amax, st);
    tend();
    launch_zero_counters(d_ctr, d_rctr, 4 * nroots, st);
    const int nl = S.nlevels;
    for (int L = 0; L < nl; L++) {
      tbeg(TC_ASM);
      launch_assemble<T>(d_asm + asm_l[L].off, asm_l[L].cnt, ds, F, st);
      launch_scatter<T>(d_aidx + scat_l[L].off, d_apos + scat_l[L].off, scat_l[L].cnt, d_vals, F, st);
      tend();
      for (size_t p = 0; p < steps[L].size(); p++) {
        const Step& sp = steps[L][p];
        if (sp.look.cnt) {
          tbeg(TC_UPD);
          launch_panel_update<T>(d_upd + sp.look.off, sp.look.cnt, sp.pfirst, ds, F, W, st);
          tend();
        }
        tbeg(TC_DIAG);
        launch_panel_diag<T>(d_diag + sp.diag.off, sp.diag.cnt, (int)p, ds, F, F + front_elems,
                             F + front_elems + n, pert, d_amax, d_ctr, d_rctr, st);
        tend();
        tbeg(TC_TRSM);
        launch_panel_trsm<T>(d_trsm + sp.trsm.off, sp.trsm.cnt, (int)p, sp.slot, ds, F, W, st);
        tend();
        if (sp.upd.cnt) {
          tbeg(TC_UPD);
          launch_panel_update<T>(d_upd + sp.upd.off, sp.upd.cnt, sp.pfirst, ds, F, W, st);
          tend();
        }
        if (tm_on) tm_upd_flops += sp.upd_flops;
      }
    }
    // pivot-block inverses for the solves (independent across supernodes)
    for (size_t p = 1; p < xrow_l.size(); p++) {
      tbeg(TC_DIAG);
      launch_panel_xrow<T>(d_xrow + xrow_l[p].off, xrow_l[p].cnt, (int)p, ds, F, st);
      tend();
    }
    // solve form: Z^T = (L21 X)^T beside the pivot block, then G = [X; Z] below
    tbeg(TC_TRSM);
    launch_zt<T>(d_zt, zt_all.cnt, ds, F, st);
    launch_gform<T>(d_gf, gf_all.cnt, ds, F, st);
    // the triangular sweeps' operands, one contiguous range per tree level
    launch_pack_solve<T>(d_pack, pack_cnt, ds, F, pk_fwd(PK), pk_bwd(PK), st);
    tend();
  }

  int factor_device(const double* dv, bool from_host_copy) {
    if (state < ANALYZED) return fail(KKT_ERR_STATE, "factor called before analyze");
    fallback_active = false;
    if (!from_host_copy && dv != d_vals) {
      if (hipMemcpyAsync(d_vals, dv, (size_t)nnz * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(KKT_ERR_COPY, "hipMemcpyAsync (vals D2D) failed");
    }
    if (!delay_known) {
      // first factor after a pattern-only analysis: look at the diagonal once
      std::vector<double> hv(nnz);
      if (hipMemcpyAsync(hv.data(), d_vals, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return fail(KKT_ERR_COPY, "hipMemcpy (vals D2H) failed");
      std::vector<unsigned char> mask;
      delay_known = true;
      if (make_mask(hv.data(), mask)) {
        int rc = analyze_with_mask(mask.data());
        if (rc) return rc;
        DevOpLock lk(g_capture_mu);
        if (hipMemcpy(d_vals, hv.data(), (size_t)nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
          return fail(KKT_ERR_COPY, "hipMemcpy (vals) failed");
      }
    }
    return do_factor(prec);
  }

  int do_factor(int p) {
    int rc = ensure_numeric(p);
    if (rc) return rc;
    double t0 = g_timing ? now_ms() : 0;
    auto body = [&]() {
      launch_expand_vals(nnzf, d_map_f, d_vals, d_valsf, st);
      if (p == KKT_PREC_FP64) {
        run_factor<double>(d_F64, d_W64, d_PK64, 1e-8);
      } else {
        run_factor<float>(d_F32, d_W32, d_PK32, 1e-6);
        if (half_on && d_PK16) launch_pack_half((long long)pk16_elems(), d_PK32, d_PK16, d_ctr, st);
      }
    };
    // the factor's launch sequence is fixed by the analysis: one graph per precision,
    // unless factor kernels are being timed with events
    const int fcls = (1 << TC_ASM) | (1 << TC_DIAG) | (1 << TC_TRSM) | (1 << TC_UPD) | (1 << TC_VEC);
    const int q = (p == KKT_PREC_FP64) ? 0 : 1;
    if (graphs_on && !(tm_on && (tm_mask & fcls))) {
      if (!fexec[q] && !capture(fexec[q], body)) graphs_on = false;
    }
    if (graphs_on && fexec[q] && !(tm_on && (tm_mask & fcls))) (void)hipGraphLaunch(fexec[q], st);
    else body();
    hipError_t e = hipMemcpyAsync(&ctr, d_ctr, sizeof(FactorCounters), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return fail(KKT_ERR_COPY, hip_msg("hipMemcpyAsync (counters) failed", e));
    h_rctr.assign(4 * (size_t)nroots, 0);
    if (nroots > 0 &&
        (e = hipMemcpyAsync(h_rctr.data(), d_rctr, 4 * (size_t)nroots * sizeof(int), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
      return fail(KKT_ERR_COPY, hip_msg("hipMemcpyAsync (component counters) failed", e));
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
      return fail(KKT_ERR_FACTORIZATION, hip_msg("factor kernels failed", e));
    tflush();
    if (hipGetLastError() != hipSuccess) return fail(KKT_ERR_FACTORIZATION, "kernel launch failed");
    fprec = p;
    have16 = (p == KKT_PREC_FP32_REFINE) && half_on && d_PK16 && !ctr.half_overflow;
    state = FACTORED;
    if (g_timing)
      fprintf(stderr, "[kkt-timing] factor(%s) n=%d: %.3f ms (inertia +%d/-%d, perturbed %d)\n",
              p == KKT_PREC_FP64 ? "fp64" : "fp32", n, now_ms() - t0, ctr.npos, ctr.nneg, ctr.npert);
    return KKT_OK;
  }

  // ------------------------------------------------------------------
  // BiCGStab's search-direction update p = r + beta (p - omega v), run by a
  // solve's gather when p is the right-hand side (p updated in place)
  struct BicgPre {
    const double* sc;
    const double* r;
    const double* v;
    double* p;
  };
  // F: the fronts (for D^-1), PK: the packed arrays (their update slots), PG:
  // the packed rows and columns the sweeps read (PK itself, or its half copy)
  template <typename T, typename TS>
  void run_tri_solve(const T* F, T* PK, const TS* PG, const double* d_b, double* d_x, void* svec = nullptr,
                     const BicgPre* pre = nullptr) {
    T* sv = (T*)(svec ? svec : d_svec);
    T *bp = sv, *fp = sv + n, *y = sv + 2 * (size_t)n, *xp = sv + 3 * (size_t)n, *wv = sv + 4 * (size_t)n;
    T* u = pk_slot(PK);
    const T* E0 = F + front_elems;
    const T* E1 = E0 + n;
    tbeg(TC_VEC);
    if (pre) launch_gather_bicg_p<T>(n, d_perm, pre->sc, pre->r, pre->v, pre->p, bp, st);
    else launch_gather_perm<double, T>(n, d_perm, d_b, bp, st);
    tend();
    const int nl = S.nlevels;
    // timed per launch: class TC_SLEVEL = forward launches, TC_SGEMV = backward
    // class TC_SLEVEL events bracket each k_fwd_level launch on its own
    for (int L = 0; L < nl; L++) {
      if (!ff_l[L].cnt) continue;
      if (lev_timing) (void)hipEventRecord(lev_ev[4 * L], st);
      tbeg(TC_SLEVEL);
      launch_fwd_level<T>(d_ff + ff_l[L].off, ff_l[L].cnt, ds, pk_fwd(PG), bp, fp, y, u, st);
      tend();
      if (fs_l[L].cnt) {
        tbeg(TC_SLEVEL);
        launch_fwd_level<T>(d_ff + fs_l[L].off, fs_l[L].cnt, ds, pk_fwd(PG), bp, fp, y, u, st);
        tend();
      }
      if (lev_timing) (void)hipEventRecord(lev_ev[4 * L + 1], st);
    }
    tbeg(TC_VEC);
    launch_dsolve<T>(n, E0, E1, d_piv, ds.wpos, y, wv, st);
    tend();
    for (int L = nl - 1; L >= 0; L--) {
      if (!bc_l[L].cnt) continue;
      if (lev_timing) (void)hipEventRecord(lev_ev[4 * L + 2], st);
      tbeg(TC_SGEMV);
      launch_bwd_cols<T>(d_bc + bc_l[L].off, bc_l[L].cnt, ds, pk_bwd(PG), wv, xp, st);
      tend();
      if (lev_timing) (void)hipEventRecord(lev_ev[4 * L + 3], st);
    }
    tbeg(TC_VEC);
    launch_scatter_perm<T, double>(n, d_perm, xp, d_x, st);
    tend();
  }
  // ---- multi-RHS solves (CuDssSolver.java:814-826, cudssBridge.cc:1486-1511;
  // the LCP block solves of KKTSolver.java:1537-1588) ----
  long long m_ldn() const { return (long long)n + 16; }
  int ensure_multi() {
    const size_t esz = std::max(sizeof(double), sizeof(float));
    const size_t need = (size_t)SOLVE_NRHS * (4 * (size_t)m_ldn() + (size_t)u_elems + 16 + (size_t)v_elems + 16) * esz;
    const size_t needv = ((size_t)2 * SOLVE_NRHS + 4 * (size_t)MULTI_BLOCK) * (size_t)n * sizeof(double);
    if (msvec_bytes >= need && mvec_bytes >= needv && d_mdot) return KKT_OK;
    DevOpLock lk(g_capture_mu);
    drop_graphs();
    if (d_msvec) (void)hipFree(d_msvec);
    if (d_mvec) (void)hipFree(d_mvec);
    if (d_mdot) (void)hipFree(d_mdot);
    if (h_mdot) (void)hipHostFree(h_mdot);
    if (d_mpart) (void)hipFree(d_mpart);
    d_msvec = nullptr;
    d_mvec = nullptr;
    d_mdot = h_mdot = d_mpart = nullptr;
    msvec_bytes = mvec_bytes = 0;
    if (hipMalloc(&d_msvec, need) != hipSuccess || hipMalloc((void**)&d_mvec, needv) != hipSuccess ||
        hipMalloc((void**)&d_mdot, MULTI_BLOCK * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&d_mpart, dot_m_partials() * sizeof(double)) != hipSuccess ||
        hipHostMalloc((void**)&h_mdot, MULTI_BLOCK * sizeof(double), hipHostMallocDefault) != hipSuccess)
      return fail(KKT_ERR_ALLOC, "hipMalloc failed (multi-RHS solve)");
    msvec_bytes = need;
    mvec_bytes = needv;
    return KKT_OK;
  }
  double* m_gin() { return d_mvec; }
  double* m_gout() { return d_mvec + (size_t)SOLVE_NRHS * n; }
  // i = 0 residuals, 1 corrections, 2 host-path right-hand sides, 3 host-path solutions
  double* m_blk(int i) { return d_mvec + ((size_t)2 * SOLVE_NRHS + (size_t)i * MULTI_BLOCK) * n; }
  // nr right-hand sides B (column j at B + j ldb) to X, one launch sequence
  template <typename T>
  void run_tri_solve_multi(const T* F, const double* B, long long ldb, double* X, long long ldx, int nr) {
    const long long ldn = m_ldn(), ldu = u_elems + 16, ldw = v_elems + 16;
    T* sv = (T*)d_msvec;
    T *bp = sv, *fp = bp + SOLVE_NRHS * ldn, *y = fp + SOLVE_NRHS * ldn, *xp = y + SOLVE_NRHS * ldn,
      *u = xp + SOLVE_NRHS * ldn, *wv = u + SOLVE_NRHS * ldu;
    const T* E0 = F + front_elems;
    const T* E1 = E0 + n;
    const int nl = S.nlevels;
    tbeg(TC_VEC);
    launch_gather_perm_m<double, T>(n, nr, d_perm, B, ldb, bp, ldn, st);
    tend();
    for (int L = 0; L < nl; L++) {
      tbeg(TC_SLEVEL);
      launch_fwd_gather_m<T>(d_fg + fg_l[L].off, fg_l[L].cnt, ds, F, nr, bp, fp, y, u, ldn, ldu, st);
      launch_fwd_rows_m<T>(d_fr + fr_l[L].off, fr_l[L].cnt, ds, F, nr, fp, y, u, ldn, ldu, st);
      tend();
    }
    tbeg(TC_SLEVEL);
    launch_dsolve_m<T>(n, nr, E0, E1, d_piv, ds.wpos, y, wv, ldn, ldw, st);
    tend();
    for (int L = nl - 1; L >= 0; L--) {
      tbeg(TC_SGEMV);
      launch_bwd_cols_m<T>(d_bc + bc_l[L].off, bc_l[L].cnt, ds, F, nr, wv, xp, ldn, ldw, st);
      tend();
    }
    tbeg(TC_VEC);
    launch_scatter_perm_m<T, double>(n, nr, d_perm, xp, ldn, X, ldx, st);
    tend();
  }
  void tri_solve_multi_eager(const double* B, long long ldb, double* X, long long ldx, int nr) {
    if (fprec == KKT_PREC_FP64) run_tri_solve_multi<double>(d_F64, B, ldb, X, ldx, nr);
    else run_tri_solve_multi<float>(d_F32, B, ldb, X, ldx, nr);
  }
  // captured once per precision and count, reading set 0 and writing set 1
  // (ld = n); other buffers are copied in and out
  void tri_solve_multi(const double* B, double* X, int nr) {
    const int solve_cls = (1 << TC_SLEVEL) | (1 << TC_SGEMV) | (1 << TC_VEC);
    if (!graphs_on || (tm_on && (tm_mask & solve_cls))) { tri_solve_multi_eager(B, n, X, n, nr); return; }
    const int q = (fprec == KKT_PREC_FP64) ? 0 : 1;
    double* gin = m_gin();
    double* gout = m_gout();
    if (!mexec[q][nr - 1] && !capture(mexec[q][nr - 1], [&]() { tri_solve_multi_eager(gin, n, gout, n, nr); })) {
      graphs_on = false;
      tri_solve_multi_eager(B, n, X, n, nr);
      return;
    }
    const size_t bytes = (size_t)nr * n * sizeof(double);
    if (B != gin) (void)hipMemcpyAsync(gin, B, bytes, hipMemcpyDeviceToDevice, st);
    (void)hipGraphLaunch(mexec[q][nr - 1], st);
    if (X != gout) (void)hipMemcpyAsync(X, gout, bytes, hipMemcpyDeviceToDevice, st);
  }
  // squared norms of nb columns (ld = n) into out (one synchronization), each as
  // launch_dot computes it
  int norms_multi(const double* A, int nb, double* out) {
    tbeg(TC_VEC);
    for (int j0 = 0; j0 < nb; j0 += SOLVE_NRHS) {
      const double* a = A + (size_t)j0 * n;
      launch_dot_m(n, std::min(SOLVE_NRHS, nb - j0), a, a, n, d_mpart, d_mdot + j0, st);
    }
    tend();
    const int nr = nb;
    if (hipMemcpyAsync(h_mdot, d_mdot, nr * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(KKT_ERR_REDUCTION, "reduction failed");
    for (int j = 0; j < nr; j++) out[j] = h_mdot[j];
    return KKT_OK;
  }
  // refined solves of nrhs columns (ld = n) on device arrays, MULTI_BLOCK columns
  // at a time, whose triangular solves go SOLVE_NRHS at a time.  Per column the
  // same steps as solve_device: FP64 factors take ir_steps refinement steps, FP32
  // factors refine to refine_tol; a column's correction is applied only while it
  // is still refining, so each result equals its single-RHS solve.  The columns
  // of a block refine in lockstep (one host synchronization per step).  A
  // stalled column falls back as solve_device does (FP64 refactor; then the
  // block's columns are solved one by one).
  int solve_multi_device(int nrhs, const double* d_B, double* d_X) {
    if (state != FACTORED) return fail(KKT_ERR_STATE, "solve called before factor");
    int rc = ensure_multi();
    if (rc) return rc;
    double* R = m_blk(0);
    double* D = m_blk(1);
    for (int c0 = 0; c0 < nrhs; c0 += MULTI_BLOCK) {
      const int nb = std::min(MULTI_BLOCK, nrhs - c0);
      const double* B = d_B + (size_t)c0 * n;
      double* X = d_X + (size_t)c0 * n;
      for (int j0 = 0; j0 < nb; j0 += SOLVE_NRHS)
        tri_solve_multi(B + (size_t)j0 * n, X + (size_t)j0 * n, std::min(SOLVE_NRHS, nb - j0));
      const bool to_tol = fprec == KKT_PREC_FP32_REFINE || fallback_active;
      const int want = to_tol ? refine_max : ir_steps;
      last_refine_steps = 0;
      if (want <= 0) continue;
      double bn[MULTI_BLOCK], prev[MULTI_BLOCK], rn2[MULTI_BLOCK];
      int stall[MULTI_BLOCK];
      bool active[MULTI_BLOCK];
      if ((rc = norms_multi(B, nb, bn))) return rc;
      for (int j = 0; j < nb; j++) {
        bn[j] = std::sqrt(bn[j]);
        if (bn[j] == 0.0) bn[j] = 1.0;
        prev[j] = 0.0;
        stall[j] = 0;
        active[j] = true;
      }
      bool fallback = false;
      for (int step = 0;; step++) {
        // residuals of every column (those already done are not used)
        tbeg(TC_SPMV);
        for (int j0 = 0; j0 < nb; j0 += SOLVE_NRHS)
          launch_spmv_resid_m(n, std::min(SOLVE_NRHS, nb - j0), d_rowptr_f, d_col_f, d_valsf, X + (size_t)j0 * n,
                              B + (size_t)j0 * n, R + (size_t)j0 * n, n, st);
        tend();
        if ((rc = norms_multi(R, nb, rn2))) return rc;
        int nact = 0;
        for (int j = 0; j < nb; j++) {
          if (!active[j]) continue;
          const double rel = std::sqrt(rn2[j]) / bn[j];
          last_relres = rel;
          if (to_tol) {
            if (rel <= refine_tol) { active[j] = false; continue; }
            if (step > 0 && rel > 0.5 * prev[j]) stall[j]++; else stall[j] = 0;
            if (step >= refine_max || stall[j] >= 2 || !std::isfinite(rel)) fallback = true;
          } else if (step >= ir_steps) {
            active[j] = false;
            continue;
          }
          prev[j] = rel;
          nact++;
        }
        if (fallback || nact == 0) break;
        // corrections for the groups that hold a refining column
        for (int j0 = 0; j0 < nb; j0 += SOLVE_NRHS) {
          const int nr = std::min(SOLVE_NRHS, nb - j0);
          bool any = false;
          for (int j = j0; j < j0 + nr; j++) any = any || active[j];
          if (any) tri_solve_multi(R + (size_t)j0 * n, D + (size_t)j0 * n, nr);
        }
        tbeg(TC_VEC);
        for (int j = 0; j < nb; j++)
          if (active[j]) launch_axpby(n, 1.0, D + (size_t)j * n, 1.0, X + (size_t)j * n, st);
        tend();
        last_refine_steps++;
      }
      if (fallback) {
        // as solve_device: the FP64 refactor with the same values (once), then
        // each column of the block solved and refined to refine_tol on it; a
        // column that still stalls returns KKT_ERR_REFINE_STALL
        if (fprec == KKT_PREC_FP32_REFINE && (rc = fallback_to_fp64())) return rc;
        for (int j = 0; j < nb; j++)
          if ((rc = solve_device(B + (size_t)j * n, X + (size_t)j * n))) return rc;
      }
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_SOLVE, "solve kernels failed");
    tflush();
    return KKT_OK;
  }
  // ---- solve-shape autotuner ----
  // Candidates: the nine uniform (forward, backward) elements-per-lane pairs from
  // {8, 16, 32}, and a per-level pick from timing every level's launches eagerly
  // (HIP events, best of `reps`) for 4, 8, 16, 32.  Each candidate is then timed
  // the way solves run (its captured graph, or eager launches when graphs are
  // off) over 4 reps triangular solves; the fastest whole solve wins, so the
  // default (16, 16) is kept unless another candidate beats it.  An FP32 factor
  // with a half-precision copy is tuned on the copy: the iterative solve's
  // preconditioner applications are most of its solves.
  bool lev_timing = false;
  std::vector<hipEvent_t> lev_ev;
  int time_solves(int nrep, double* b, double* x, float* ms, bool half) {
    tri_solve(b, x, half);  // captures the graph of the current items
    hipEvent_t e0 = lev_ev[0], e1 = lev_ev[1];
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < nrep; r++) tri_solve(b, x, half);
    (void)hipEventRecord(e1, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_DEVICE, "autotune solve failed");
    if (hipEventElapsedTime(ms, e0, e1) != hipSuccess) return fail(KKT_ERR_DEVICE, "autotune timing failed");
    *ms /= (float)nrep;
    return KKT_OK;
  }
  int autotune_solve(int reps) {
    if (state < FACTORED) return fail(KKT_ERR_STATE, "factor() not previously called");
    reps = std::max(1, std::min(reps, 20));
    const int nl = (int)sv_lv.size();
    if ((int)lev_ev.size() < std::max(4 * nl, 2)) {
      DevOpLock lk(g_capture_mu);
      while ((int)lev_ev.size() < std::max(4 * nl, 2)) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return fail(KKT_ERR_DEVICE, "hipEventCreate failed");
        lev_ev.push_back(e);
      }
    }
    double* b = d_vec + 12 * (size_t)n;
    double* x = d_vec + 13 * (size_t)n;
    launch_fill(n, 1.0, b, st);
    const bool half = use16();
    int rc;
    // per-level pick from eager per-launch timing: the lane shape of both sweeps
    // and the one- or two-launch forward form
    // forms per lane shape: one launch with 1, 2, 4 or 8 row groups per block,
    // or two launches
    static const int cand[] = {4, 8, 16, 32};
    static const int rgs[] = {1, 2, 4, 8, 0};  // 0: the two-launch form
    std::vector<float> bfr(nl, 1e30f), bbc(nl, 1e30f);
    std::vector<int> cfr(nl, 16), cbc(nl, 16), cfu(nl, 1), crg(nl, 1);
    for (int pass = 0; pass < 20; pass++) {
      const int c = cand[pass / 5], rgp = rgs[pass % 5], fu = rgp ? 1 : 0;
      fr_epl.assign(nl, c);
      bc_epl.assign(nl, c);
      fr_fuse.assign(nl, fu);
      fr_rg.assign(nl, rgp ? rgp : 1);
      if ((rc = build_solve_items())) return rc;
      tri_solve_eager(b, x, nullptr, half);
      std::vector<float> tfr(nl, 1e30f), tbc(nl, 1e30f);
      for (int r = 0; r < reps; r++) {
        lev_timing = true;
        tri_solve_eager(b, x, nullptr, half);
        lev_timing = false;
        if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_DEVICE, "autotune solve failed");
        for (int L = 0; L < nl; L++) {
          float ms;
          if (ff_l[L].cnt && hipEventElapsedTime(&ms, lev_ev[4 * L], lev_ev[4 * L + 1]) == hipSuccess)
            tfr[L] = std::min(tfr[L], ms);
          if (bc_l[L].cnt && hipEventElapsedTime(&ms, lev_ev[4 * L + 2], lev_ev[4 * L + 3]) == hipSuccess)
            tbc[L] = std::min(tbc[L], ms);
        }
      }
      for (int L = 0; L < nl; L++) {
        if (tfr[L] < bfr[L]) { bfr[L] = tfr[L]; cfr[L] = c; cfu[L] = fu; crg[L] = rgp ? rgp : 1; }
        if (rgp == 1 && tbc[L] < bbc[L]) { bbc[L] = tbc[L]; cbc[L] = c; }
      }
    }
    // whole-solve timing of every candidate: the nine uniform shape pairs with
    // one forward launch per level (default row groups), and the per-level picks
    struct Cand { std::vector<int> f, b, fu, rg; };
    std::vector<Cand> cands;
    const std::vector<int> all1(nl, 1);
    cands.push_back({std::vector<int>(nl, 16), std::vector<int>(nl, 16), all1, default_row_groups(16)});
    for (int f : {8, 16, 32})
      for (int g : {8, 16, 32})
        if (f != 16 || g != 16)
          cands.push_back({std::vector<int>(nl, f), std::vector<int>(nl, g), all1, default_row_groups(f)});
    cands.push_back({cfr, cbc, cfu, crg});
    float best = 1e30f;
    size_t bi = 0;
    for (size_t q = 0; q < cands.size(); q++) {
      fr_epl = cands[q].f;
      bc_epl = cands[q].b;
      fr_fuse = cands[q].fu;
      fr_rg = cands[q].rg;
      if ((rc = build_solve_items())) return rc;
      drop_graphs();
      float ms;
      if ((rc = time_solves(4 * reps, b, x, &ms, half))) return rc;
      if (ms < best) { best = ms; bi = q; }
      if (g_timing)
        fprintf(stderr, "[kkt-timing] autotune(%s) candidate %zu: %.3f ms per triangular solve\n",
                fprec == KKT_PREC_FP64 ? "fp64" : "fp32", q, ms);
    }
    fr_epl = cands[bi].f;
    bc_epl = cands[bi].b;
    fr_fuse = cands[bi].fu;
    fr_rg = cands[bi].rg;
    if (g_timing) {
      fprintf(stderr, "[kkt-timing] autotune(%s): candidate %zu, two-launch forward levels:",
              fprec == KKT_PREC_FP64 ? "fp64" : "fp32", bi);
      for (int L = 0; L < nl; L++)
        if (!fr_fuse[L]) fprintf(stderr, " %d", L);
      fprintf(stderr, "\n");
    }
    if ((rc = build_solve_items())) return rc;
    drop_graphs();
    tflush();
    return KKT_OK;
  }
  // half: the iterative solve's preconditioner (the half-precision copy, when
  // the FP32 factor has one)
  void tri_solve_eager(const double* d_b, double* d_x, const BicgPre* pre = nullptr, bool half = false) {
    if (fprec == KKT_PREC_FP64) run_tri_solve<double>(d_F64, d_PK64, (const double*)d_PK64, d_b, d_x, nullptr, pre);
    else if (half && use16()) run_tri_solve<float>(d_F32, d_PK32, (const _Float16*)d_PK16, d_b, d_x, nullptr, pre);
    else run_tri_solve<float>(d_F32, d_PK32, (const float*)d_PK32, d_b, d_x, nullptr, pre);
  }
  // The launch sequence of a solve is fixed by the analysis, so it is captured
  // once into a HIP graph (reading vec(12), writing vec(13)) and replayed.
  int solve_q(bool half) const { return fprec == KKT_PREC_FP64 ? 0 : (half && use16() ? 2 : 1); }
  void tri_solve(const double* d_b, double* d_x, bool half = false) {
    const int solve_cls = (1 << TC_SLEVEL) | (1 << TC_SGEMV) | (1 << TC_VEC);
    if (!graphs_on || (tm_on && (tm_mask & solve_cls))) { tri_solve_eager(d_b, d_x, nullptr, half); return; }
    const int q = solve_q(half);
    double* gin = vec(12);
    double* gout = vec(13);
    if (!gexec[q] && !capture(gexec[q], [&]() { tri_solve_eager(gin, gout, nullptr, half); })) {
      graphs_on = false;
      tri_solve_eager(d_b, d_x, nullptr, half);
      return;
    }
    if (d_b != gin) (void)hipMemcpyAsync(gin, d_b, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st);
    (void)hipGraphLaunch(gexec[q], st);
    if (d_x != gout) (void)hipMemcpyAsync(d_x, gout, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st);
  }

  double* vec(int i) { return d_vec + (size_t)i * n; }

  int dotv(const double* a, const double* b, double& out) {
    tbeg(TC_VEC);
    launch_dot(n, a, b, d_partial, d_scal, st);
    tend();
    if (hipMemcpyAsync(&out, d_scal, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(KKT_ERR_REDUCTION, "reduction failed");
    return KKT_OK;
  }
  void spmv(const double* x, const double* b, double* y, int mode) {
    tbeg(TC_SPMV);
    launch_spmv(n, d_rowptr_f, d_col_f, d_valsf, x, b, y, mode, st);
    tend();
  }

  // FP64 refactor with the current values after FP32 refinement stalled; the
  // handle keeps the FP64 factor (and refines to refine_tol) until the next
  // factor call
  int fallback_to_fp64() {
    fp64_fallbacks++;
    int rc = do_factor(KKT_PREC_FP64);
    if (rc) return rc;
    fallback_active = true;
    return KKT_OK;
  }

  // Refined solve on device vectors; d_b and d_x must not alias.  FP32 factors
  // (and the FP64 factor of a fallback) refine to refine_tol: a stall (no 2x
  // contraction for two steps, more than refine_max steps, or a non-finite
  // residual) on an FP32 factor triggers the FP64 refactor and the refinement
  // starts again from the FP64 solve; a stall on the FP64 factor returns
  // KKT_ERR_REFINE_STALL with the best solution left in d_x.  Plain FP64
  // factors take ir_steps refinement steps, as the bridge's IR setting.
  int solve_device(const double* d_b, double* d_x) {
    if (state != FACTORED) return fail(KKT_ERR_STATE, "solve called before factor");
    double t0 = g_timing ? now_ms() : 0;
    tri_solve(d_b, d_x);
    last_refine_steps = 0;
    const bool to_tol = fprec == KKT_PREC_FP32_REFINE || fallback_active;
    const int want = to_tol ? refine_max : ir_steps;
    if (want > 0) {
      double* r = vec(0);
      double* dx = vec(1);
      double bn2;
      int rc = dotv(d_b, d_b, bn2);
      if (rc) return rc;
      double bn = std::sqrt(bn2);
      if (bn == 0.0) bn = 1.0;
      double prev = 0.0;
      int stall = 0;
      for (int step = 0;; step++) {
        spmv(d_x, d_b, r, 1);
        double rn2;
        if ((rc = dotv(r, r, rn2))) return rc;
        double rel = std::sqrt(rn2) / bn;
        last_relres = rel;
        if (to_tol) {
          if (rel <= refine_tol) break;
          if (step > 0 && rel > 0.5 * prev) stall++; else stall = 0;
          if (step >= refine_max || stall >= 2 || !std::isfinite(rel)) {
            if (fprec == KKT_PREC_FP32_REFINE) {
              if ((rc = fallback_to_fp64())) return rc;
              tri_solve(d_b, d_x);
              step = -1;
              stall = 0;
              prev = 0.0;
              continue;
            }
            (void)hipStreamSynchronize(st);
            tflush();
            char msg[160];
            snprintf(msg, sizeof msg, "refinement stalled at relative residual %.3g > %.3g after the FP64 refactor",
                     rel, refine_tol);
            return fail(KKT_ERR_REFINE_STALL, msg);
          }
        } else if (step >= ir_steps) {
          break;
        }
        prev = rel;
        tri_solve(r, dx);
        tbeg(TC_VEC);
        launch_axpby(n, 1.0, dx, 1.0, d_x, st);
        tend();
        last_refine_steps++;
      }
    }
    if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_SOLVE, "solve kernels failed");
    tflush();
    if (g_timing)
      fprintf(stderr, "[kkt-timing] solve n=%d: %.3f ms (refine steps %d, rel resid %.3g)\n", n,
              now_ms() - t0, last_refine_steps, last_relres);
    return KKT_OK;
  }

  // ---- refined solve without allocation (kkt_solve_device_noalloc) ----
  // The per-call working set -- the triangular solves' vectors (permuted
  // right-hand side, forward and D^-1 vectors, update and backward vectors, in
  // the factor precision), the refinement residual and correction, the dot
  // products' partial sums and result -- lives in the caller-supplied scratch
  // region `scr` of `cap` bytes instead of the handle's buffers.  Launches are
  // eager (no graph capture or instantiation), no timing events are created,
  // and a stalled FP32 refinement returns KKT_ERR_REFINE_STALL instead of
  // refactoring in FP64 (which would allocate the FP64 fronts).
  static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
  size_t noalloc_bytes() const {
    const size_t esz = (fprec == KKT_PREC_FP64) ? sizeof(double) : sizeof(float);
    const size_t sv = align256(svec_elems() * esz);
    return sv + 2 * align256((size_t)n * sizeof(double)) + align256(1024 * sizeof(double)) + 256;
  }
  int solve_device_noalloc(const double* d_b, double* d_x, unsigned char* scr, size_t cap) {
    if (state != FACTORED) return fail(KKT_ERR_STATE, "solve called before factor");
    const size_t need = noalloc_bytes();
    if (need > cap) {
      char msg[256];
      snprintf(msg, sizeof msg, "this system's solve needs %zu bytes of scratch, more than the %zu bytes the "
               "no-allocation solve holds (use kkt_solve_device)", need, cap);
      return fail(KKT_ERR_ARGUMENT, msg);
    }
    const size_t esz = (fprec == KKT_PREC_FP64) ? sizeof(double) : sizeof(float);
    unsigned char* p = scr;
    void* svec = p;
    p += align256(svec_elems() * esz);
    double* r = (double*)p;
    p += align256((size_t)n * sizeof(double));
    double* dx = (double*)p;
    p += align256((size_t)n * sizeof(double));
    double* part = (double*)p;
    p += align256(1024 * sizeof(double));
    double* scal = (double*)p;
    auto tri = [&](const double* b, double* x) {
      if (fprec == KKT_PREC_FP64) run_tri_solve<double>(d_F64, d_PK64, (const double*)d_PK64, b, x, svec);
      else run_tri_solve<float>(d_F32, d_PK32, (const float*)d_PK32, b, x, svec);
    };
    auto dot = [&](const double* a, const double* b, double& out) -> int {
      launch_dot(n, a, b, part, scal, st);
      if (hipMemcpyAsync(&out, scal, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return fail(KKT_ERR_REDUCTION, "reduction failed");
      return KKT_OK;
    };
    const bool tm = tm_on;
    tm_on = false;  // no timing events (creating them would allocate)
    tri(d_b, d_x);
    last_refine_steps = 0;
    int rc = KKT_OK;
    const bool to_tol = fprec == KKT_PREC_FP32_REFINE || fallback_active;
    const int want = to_tol ? refine_max : ir_steps;
    if (want > 0) {
      double bn2;
      if ((rc = dot(d_b, d_b, bn2))) { tm_on = tm; return rc; }
      double bn = std::sqrt(bn2);
      if (bn == 0.0) bn = 1.0;
      double prev = 0.0;
      int stall = 0;
      for (int step = 0;; step++) {
        launch_spmv(n, d_rowptr_f, d_col_f, d_valsf, d_x, d_b, r, 1, st);
        double rn2;
        if ((rc = dot(r, r, rn2))) break;
        const double rel = std::sqrt(rn2) / bn;
        last_relres = rel;
        if (to_tol) {
          if (rel <= refine_tol) break;
          if (step > 0 && rel > 0.5 * prev) stall++; else stall = 0;
          if (step >= refine_max || stall >= 2 || !std::isfinite(rel)) {
            char msg[200];
            snprintf(msg, sizeof msg, "refinement stalled at relative residual %.3g > %.3g (the no-allocation "
                     "solve does not refactor in FP64)", rel, refine_tol);
            rc = fail(KKT_ERR_REFINE_STALL, msg);
            break;
          }
        } else if (step >= ir_steps) {
          break;
        }
        prev = rel;
        tri(r, dx);
        launch_axpby(n, 1.0, dx, 1.0, d_x, st);
        last_refine_steps++;
      }
    }
    tm_on = tm;
    if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_SOLVE, "solve kernels failed");
    return rc;
  }

  // ---- per-launch kernel times without per-launch events ----
  // One forward sweep (what 0), one backward sweep (1) or one residual SpMV (2)
  // of the current factor, or a forward (3) or backward (4) sweep reading an FP32
  // factor's half-precision copy, captured into a graph and replayed `reps` times
  // between two HIP events on the handle's stream: graph-replayed kernels run
  // back to back, so elapsed / (reps x launches) is the kernels' average
  // duration, as a kernel trace reports it.  Replays are idempotent (each
  // recomputes the same vectors from the same inputs).
  int time_kernels(int what, int reps, double* ms_per_launch, int* launches) {
    if (state != FACTORED) return fail(KKT_ERR_STATE, "factor() not previously called");
    if (what < 0 || what > 4 || reps < 1) return fail(KKT_ERR_ARGUMENT, "bad kernel group or repetitions");
    if (what >= 3 && !use16()) return fail(KKT_ERR_STATE, "the factor has no half-precision solve copy");
    const bool f64 = fprec == KKT_PREC_FP64, half = what >= 3;
    if (half) what -= 3;
    void* sv = d_svec;
    const int nl = S.nlevels;
    int cnt = 0;
    for (int L = 0; L < nl; L++) {
      if (what == 0) cnt += (ff_l[L].cnt > 0) + (fs_l[L].cnt > 0);
      if (what == 1) cnt += bc_l[L].cnt > 0;
    }
    if (what == 2) cnt = 1;
    double* b = vec(12);
    double* x = vec(13);
    // inputs of the sweeps: one eager triangular solve fills every vector
    launch_fill(n, 1.0, b, st);
    tri_solve_eager(b, x);
    auto body = [&]() {
      if (what == 2) {
        launch_spmv(n, d_rowptr_f, d_col_f, d_valsf, x, b, vec(0), 1, st);
        return;
      }
      auto sweep = [&](auto* PK, const auto* PG) {
        typedef typename std::remove_pointer<decltype(PK)>::type T;
        T* svT = (T*)sv;
        T *bp = svT, *fp = svT + n, *y = svT + 2 * (size_t)n, *xp = svT + 3 * (size_t)n,
          *wv = svT + 4 * (size_t)n;
        T* u = pk_slot(PK);
        for (int L = 0; L < nl; L++) {
          if (what == 0) {
            if (ff_l[L].cnt) launch_fwd_level<T>(d_ff + ff_l[L].off, ff_l[L].cnt, ds, pk_fwd(PG), bp, fp, y, u, st);
            if (fs_l[L].cnt) launch_fwd_level<T>(d_ff + fs_l[L].off, fs_l[L].cnt, ds, pk_fwd(PG), bp, fp, y, u, st);
          } else {
            const int Lb = nl - 1 - L;
            if (bc_l[Lb].cnt) launch_bwd_cols<T>(d_bc + bc_l[Lb].off, bc_l[Lb].cnt, ds, pk_bwd(PG), wv, xp, st);
          }
        }
      };
      if (f64) sweep(d_PK64, (const double*)d_PK64);
      else if (half) sweep(d_PK32, (const _Float16*)d_PK16);
      else sweep(d_PK32, (const float*)d_PK32);
    };
    hipGraphExec_t ex = nullptr;
    if (!capture(ex, body)) return fail(KKT_ERR_DEVICE, "graph capture failed (kernel timing)");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    {
      DevOpLock lk(g_capture_mu);
      if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        (void)hipGraphExecDestroy(ex);
        return fail(KKT_ERR_DEVICE, "hipEventCreate failed");
      }
    }
    (void)hipGraphLaunch(ex, st);  // warm
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; r++) (void)hipGraphLaunch(ex, st);
    (void)hipEventRecord(e1, st);
    float ms = 0;
    hipError_t err = hipStreamSynchronize(st);
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(ex);
    if (err != hipSuccess) return fail(KKT_ERR_DEVICE, hip_msg("kernel timing failed", err));
    *ms_per_launch = cnt ? (double)ms / ((double)reps * cnt) : 0.0;
    *launches = cnt;
    return KKT_OK;
  }

  int set_values_device(const double* dv) {
    if (dv != d_vals &&
        hipMemcpyAsync(d_vals, dv, (size_t)nnz * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(KKT_ERR_COPY, "hipMemcpyAsync (vals) failed");
    launch_expand_vals(nnzf, d_map_f, d_vals, d_valsf, st);
    return KKT_OK;
  }

  // ---- device-side value assembly (CuDssSolver.java:421-788) ----
  // The plan of a slot list: its valid contributions (slot >= 0) stably sorted by
  // slot, grouped per distinct slot.  Plans are cached by the exact slot list.
  int plan_for(const int* slots, long long cnt, AddPlan*& out) {
    for (AddPlan& p : add_plans)
      if ((long long)p.slots.size() == cnt && std::equal(slots, slots + cnt, p.slots.begin())) {
        p.used = ++add_clock;
        out = &p;
        return KKT_OK;
      }
    std::vector<int> idx;
    idx.reserve((size_t)cnt);
    for (long long e = 0; e < cnt; e++) {
      const int sl = slots[e];
      if (sl < 0) continue;
      if (sl >= nnz) return fail(KKT_ERR_ARGUMENT, "slot beyond the analyzed pattern's nnz");
      idx.push_back((int)e);
    }
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return slots[a] < slots[b]; });
    std::vector<int> useg, uslot;
    for (size_t q = 0; q < idx.size(); q++)
      if (q == 0 || slots[idx[q]] != slots[idx[q - 1]]) {
        useg.push_back((int)q);
        uslot.push_back(slots[idx[q]]);
      }
    useg.push_back((int)idx.size());
    if (add_plans.size() >= 16) {
      size_t lru = 0;
      for (size_t q = 1; q < add_plans.size(); q++)
        if (add_plans[q].used < add_plans[lru].used) lru = q;
      free_plan(add_plans[lru]);
      add_plans.erase(add_plans.begin() + (long)lru);
    }
    add_plans.emplace_back();
    AddPlan& p = add_plans.back();
    p.slots.assign(slots, slots + cnt);
    p.nuniq = (int)uslot.size();
    p.used = ++add_clock;
    bool okp;
    {
      DevOpLock lk(g_capture_mu);
      okp = hipMalloc(&p.d_order, std::max<size_t>(idx.size(), 1) * sizeof(int)) == hipSuccess &&
            hipMalloc(&p.d_useg, useg.size() * sizeof(int)) == hipSuccess &&
            hipMalloc(&p.d_uslot, std::max<size_t>(uslot.size(), 1) * sizeof(int)) == hipSuccess &&
            (idx.empty() || hipMemcpy(p.d_order, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess) &&
            hipMemcpy(p.d_useg, useg.data(), useg.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
            (uslot.empty() || hipMemcpy(p.d_uslot, uslot.data(), uslot.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess);
    }
    if (!okp) {
      free_plan(p);
      add_plans.pop_back();
      return fail(KKT_ERR_ALLOC, "device allocation or copy failed (assembly plan)");
    }
    out = &p;
    return KKT_OK;
  }
  // uploads the listed host arrays into one staging buffer; returns device pointers
  int stage_inputs(std::initializer_list<std::pair<const void*, size_t>> in, std::vector<const double*>& dp) {
    size_t total = 0;
    for (auto& a : in) total += (a.second + 255) & ~(size_t)255;
    if (total > stage_cap) {
      DevOpLock lk(g_capture_mu);
      if (d_stage) (void)hipFree(d_stage);
      d_stage = nullptr;
      stage_cap = 0;
      if (hipMalloc(&d_stage, std::max<size_t>(total, 256)) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (assembly inputs)");
      stage_cap = std::max<size_t>(total, 256);
    }
    size_t off = 0;
    dp.clear();
    for (auto& a : in) {
      if (a.second && hipMemcpyAsync(d_stage + off, a.first, a.second, hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(KKT_ERR_COPY, "hipMemcpyAsync (assembly inputs) failed");
      dp.push_back((const double*)(d_stage + off));
      off += (a.second + 255) & ~(size_t)255;
    }
    return KKT_OK;
  }
  int ensure_contrib(size_t cnt) {
    if (cnt <= contrib_cap) return KKT_OK;
    DevOpLock lk(g_capture_mu);
    if (d_contrib) (void)hipFree(d_contrib);
    d_contrib = nullptr;
    contrib_cap = 0;
    if (hipMalloc(&d_contrib, std::max<size_t>(cnt, 1) * sizeof(double)) != hipSuccess)
      return fail(KKT_ERR_ALLOC, "hipMalloc failed (assembly contributions)");
    contrib_cap = std::max<size_t>(cnt, 1);
    return KKT_OK;
  }
  // common tail: plan, contributions (produce(d_contrib)), ordered scatter, sync
  template <typename Fn>
  int device_add(const int* slots, long long cnt, double scale, Fn produce) {
    if (state < ANALYZED) return fail(KKT_ERR_STATE, "analyze() not previously called");
    AddPlan* pl = nullptr;
    int rc;
    if ((rc = plan_for(slots, cnt, pl))) return rc;
    if ((rc = ensure_contrib((size_t)cnt))) return rc;
    if ((rc = produce(d_contrib))) return rc;
    launch_seg_scatter(pl->nuniq, pl->d_useg, pl->d_uslot, pl->d_order, d_contrib, scale, d_vals, st);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_DEVICE, "device value assembly failed");
    return KKT_OK;
  }

  // reads `cnt` device scalars starting at slot `first` (one stream sync)
  int read_scalars(int first, int cnt, double* out) {
    if (hipMemcpyAsync(out, d_scal + first, cnt * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(KKT_ERR_REDUCTION, "reduction failed");
    return KKT_OK;
  }
  void dot_to(const double* a, const double* b, int slot) {
    tbeg(TC_VEC);
    launch_dot(n, a, b, d_partial, d_scal + slot, st);
    tend();
  }

  // BiCGStab preconditioned by the current (stale) factor: cudssBridge.cc:1705-2006
  // (x0 = M^-1 b, r_hat = r0, early exit on ||s||, breakdown threshold 1e-30).
  // Scalars stay on the device; the host reads them twice per iteration for the
  // convergence and breakdown tests.
  int bicgstab(const double* d_b, double* d_xout, double tolRel, int maxIter, int* iters) {
    if (state != FACTORED) return fail(KKT_ERR_STATE, "iterativeSolve called before factor");
    double t0 = g_timing ? now_ms() : 0;
    double *r = vec(0), *rh = vec(1), *p = vec(2), *v = vec(3), *s = vec(4), *t = vec(5),
           *y = vec(6), *z = vec(7), *x = vec(8);
    double* sc = d_scal;
    // the preconditioner M^-1 reads the half-precision copy of an FP32 factor's
    // solve arrays when there is one (products and sums stay FP32)
    tri_solve(d_b, x, true);
    spmv(x, d_b, r, 1);
    dot_to(d_b, d_b, 9);
    dot_to(r, r, 8);
    double h2[2];
    int rc;
    if ((rc = read_scalars(8, 2, h2))) return rc;
    double bn = std::sqrt(h2[1]), rn = std::sqrt(h2[0]);
    if (bn == 0.0) bn = 1.0;
    int it = 1;
    if (rn / bn > tolRel) {
      (void)hipMemcpyAsync(rh, r, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st);
      launch_fill(n, 0.0, p, st);
      launch_fill(n, 0.0, v, st);
      launch_bicg_init(sc, st);
      const double eps = 1e-30;
      // Each half-iteration ends with its scalars copied to pinned host memory;
      // both halves are captured once into HIP graphs (fixed buffers), so an
      // iteration costs two graph launches and two host synchronizations.
      // Dot products are fused into the kernels that produce their operands
      // (per-block partials, then one fixed-order reduction of one or two sums);
      // rho of the next iteration comes with the x / r update.
      double *pa = d_part2, *pb = d_part2 + std::max(spmv_blocks(n), 1);
      const int nbs = spmv_blocks(n), nbv = vec_blocks(n);
      // The scalars of an iteration (alpha, omega, rho_prev; rho = next rho) are
      // formed by the previous iteration's last reduction; p = r + beta (p - omega v)
      // runs in the gather of y = M^-1 p.
      const BicgPre pre{sc, r, v, p};
      auto head = [&]() {
        tri_solve_eager(p, y, &pre, true);            // p updated, y = M^-1 p
        tbeg(TC_SPMV);
        launch_spmv(n, d_rowptr_f, d_col_f, d_valsf, y, nullptr, v, 0, st, rh, pa, nullptr);  // v = A y
        tend();
        tbeg(TC_VEC);
        launch_reduce2(pa, nullptr, nbs, sc + 4, nullptr, st);  // (r_hat, v)
        launch_bicg_s(n, sc, r, v, s, pa, st);                   // s = r - alpha v
        launch_reduce2(pa, nullptr, nbv, sc + 5, nullptr, st);  // |s|^2
        tend();
        (void)hipMemcpyAsync(h_pin, sc, 6 * sizeof(double), hipMemcpyDeviceToHost, st);
      };
      auto tail = [&]() {
        tri_solve_eager(s, z, nullptr, true);         // z = M^-1 s
        tbeg(TC_SPMV);
        launch_spmv(n, d_rowptr_f, d_col_f, d_valsf, z, nullptr, t, 0, st, s, pa, pb);  // t = A z
        tend();
        tbeg(TC_VEC);
        launch_reduce2(pa, pb, nbs, sc + 6, sc + 7, st);  // (s, t), (t, t)
        launch_bicg_xr(n, sc, y, z, s, t, x, r, rh, pa, pb, st);  // x += alpha y + omega z; r = s - omega t
        launch_reduce2(pa, pb, nbv, sc + 8, sc + 10, st, sc);  // |r|^2, (r_hat, r); next scalars
        tend();
        (void)hipMemcpyAsync(h_pin + 6, sc + 6, 3 * sizeof(double), hipMemcpyDeviceToHost, st);
      };
      const int q = solve_q(true);
      const int cls = (1 << TC_SLEVEL) | (1 << TC_SGEMV) | (1 << TC_VEC) | (1 << TC_SPMV);
      auto run = [&](int w, auto fn) -> int {
        const bool use = graphs_on && !(tm_on && (tm_mask & cls));
        if (use && (bexec[q][w] || capture(bexec[q][w], fn))) (void)hipGraphLaunch(bexec[q][w], st);
        else fn();
        if (hipStreamSynchronize(st) != hipSuccess) return fail(KKT_ERR_REDUCTION, "reduction failed");
        return KKT_OK;
      };
      for (it = 1; it <= maxIter; it++) {
        if ((rc = run(0, head))) return rc;
        const double* h5 = h_pin;
        if (std::fabs(h5[0]) < eps) return fail(KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: rho ~ 0");
        if (std::fabs(h5[4]) < eps) return fail(KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: (r_hat,v) ~ 0");
        if (std::sqrt(h5[5]) / bn <= tolRel) {
          launch_bicg_xa(n, sc, y, x, st);            // x += alpha y
          break;
        }
        if ((rc = run(1, tail))) return rc;
        const double* h3 = h_pin + 6;
        if (h3[1] < eps) return fail(KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: (t,t) ~ 0");
        if (std::sqrt(h3[2]) / bn <= tolRel) break;
        if (std::fabs(h3[0] / h3[1]) < eps) return fail(KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: omega ~ 0");
      }
      if (it > maxIter) return fail(KKT_ERR_ITER_NO_CONVERGE, "BiCGStab did not converge within iteration limit");
    }
    if (hipMemcpyAsync(d_xout, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(KKT_ERR_SOLVE, "iterative solve failed");
    tflush();
    *iters = it;
    if (g_timing)
      fprintf(stderr, "[kkt-timing] bicgstab n=%d iters=%d: total=%.2fms (%.2fms/iter)\n", n, it,
              now_ms() - t0, (now_ms() - t0) / std::max(1, it));
    return KKT_OK;
  }
};

// =====================================================================
// Group iterative solves: the stale-factor BiCGStab of many handles (one per
// serving instance) in lockstep.  Every step of an iteration -- the gather, each
// forward tree level, D^-1, each backward level, the scatter, the SpMV, the
// reductions and the vector updates -- is one launch for all the group's active
// handles: the launch's workgroups are dealt to per-handle segments, and each
// segment runs the step's kernel body with the handle's own arguments and block
// indices (kernels.hip, group launches).  A handle's result is therefore
// bitwise the one kkt_iterative_solve_device gives it, whatever else is in the
// group, while the group pays each step's launch and latency once.  Handles
// leave the group as they converge, break down or reach maxIter.
// =====================================================================
struct kkt_group {
  hipStream_t st = nullptr;
  // device blob: per-launch segment prefixes and arguments; pinned staging
  char* d_blob = nullptr;
  char* h_blob = nullptr;
  size_t blob_bytes = 0;
  int gen = 0;             // bumped when the blob is reallocated
  double* d_sc = nullptr;  // gathered scalars: 16 per handle
  double* h_sc = nullptr;
  int sc_cap = 0;
  std::string err;
  bool has_err = false;

  int fail(int code, const std::string& m) {
    err = m;
    has_err = true;
    return code;
  }
  ~kkt_group() {
    DevOpLock lk(g_capture_mu);
    if (d_blob) (void)hipFree(d_blob);
    if (h_blob) (void)hipHostFree(h_blob);
    if (d_sc) (void)hipFree(d_sc);
    if (h_sc) (void)hipHostFree(h_sc);
    if (st) (void)hipStreamDestroy(st);
  }
  int ensure(size_t bytes, int nh) {
    DevOpLock lk(g_capture_mu);
    if (bytes > blob_bytes) {
      if (d_blob) (void)hipFree(d_blob);
      if (h_blob) (void)hipHostFree(h_blob);
      d_blob = h_blob = nullptr;
      blob_bytes = 0;
      const size_t cap = std::max(bytes + bytes / 2, (size_t)1 << 20);
      if (hipMalloc((void**)&d_blob, cap) != hipSuccess ||
          hipHostMalloc((void**)&h_blob, cap, hipHostMallocDefault) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (group tables)");
      blob_bytes = cap;
      gen++;
    }
    if (nh > sc_cap) {
      if (d_sc) (void)hipFree(d_sc);
      if (h_sc) (void)hipHostFree(h_sc);
      d_sc = h_sc = nullptr;
      sc_cap = 0;
      if (hipMalloc((void**)&d_sc, (size_t)16 * nh * sizeof(double)) != hipSuccess ||
          hipHostMalloc((void**)&h_sc, (size_t)16 * nh * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return fail(KKT_ERR_ALLOC, "hipMalloc failed (group scalars)");
      sc_cap = nh;
    }
    return KKT_OK;
  }
};

namespace {

// One group launch under construction: its family, segment prefixes and
// arguments (packed into the blob at `off`).
enum GroupFam { GF_VEC, GF_RED, GF_SPMV, GF_FWD, GF_BWD };
struct GLaunch {
  int fam = GF_VEC;
  std::vector<int> pre{0};
  std::vector<char> args;
  size_t off = 0;  // blob offset of pre, args at off + args_off
  size_t args_off = 0;
  template <typename A>
  void add(const A& a, int nblk) {
    if (nblk <= 0) return;
    const size_t o = args.size();
    args.resize(o + sizeof(A));
    std::memcpy(args.data() + o, &a, sizeof(A));
    pre.push_back(pre.back() + nblk);
  }
  int nseg() const { return (int)pre.size() - 1; }
  int nblk() const { return pre.back(); }
};

// Lays the launches out in the staging buffer (16-byte aligned sections) from
// byte `base`; returns the end.
bool segs_fit(const std::vector<GLaunch>& L) {
  for (const GLaunch& g : L)
    if (g.nseg() > SEG_MAX) return false;
  return true;
}

size_t pack_launches(std::vector<GLaunch>& L, char* h, size_t base) {
  size_t o = base;
  auto al = [](size_t v) { return (v + 15) & ~(size_t)15; };
  for (GLaunch& g : L) {
    o = al(o);
    g.off = o;
    if (h) std::memcpy(h + o, g.pre.data(), g.pre.size() * sizeof(int));
    o = al(o + g.pre.size() * sizeof(int));
    g.args_off = o;
    if (h) std::memcpy(h + o, g.args.data(), g.args.size());
    o += g.args.size();
  }
  return o;
}

template <typename T, typename S>
void run_launches(const std::vector<GLaunch>& L, const char* d, hipStream_t st) {
  for (const GLaunch& g : L) {
    // the group kernels' segment table holds SEG_MAX + 1 prefixes (checked by
    // the driver before anything is uploaded)
    if (!g.nblk() || g.nseg() > SEG_MAX) continue;
    const int* pre = (const int*)(d + g.off);
    const void* a = d + g.args_off;
    switch (g.fam) {
      case GF_VEC: launch_seg_vec<T>(pre, g.nseg(), (const SegVec*)a, g.nblk(), st); break;
      case GF_RED: launch_seg_red(pre, g.nseg(), (const SegRed*)a, g.nblk(), st); break;
      case GF_SPMV: launch_seg_spmv(pre, g.nseg(), (const SegSpmv*)a, g.nblk(), st); break;
      case GF_FWD: launch_seg_fwd<T, S>(pre, g.nseg(), (const SegFwd*)a, g.nblk(), st); break;
      case GF_BWD: launch_seg_bwd<T, S>(pre, g.nseg(), (const SegBwd*)a, g.nblk(), st); break;
    }
  }
}

SegVec vop(int op, long long n) {
  SegVec v;
  std::memset(&v, 0, sizeof v);
  v.op = op;
  v.n = n;
  return v;
}
int vblk(long long n) { return (int)((n + 255) / 256); }

}  // namespace

// the group driver, for handles of one solve precision class q (0: FP64
// arrays, 1: FP32, 2: FP32 with the half-precision preconditioner copy)
template <typename T, typename S>
static int group_bicgstab(kkt_group* g, int K, kkt_solver* const* hs, const double* const* vin,
                          const double* const* bin, double* const* xout, double tolRel, int maxIter, int* status) {
  const double eps = 1e-30;
  hipStream_t st = g->st;
  // the gathered scalars' buffer holds every handle's 16 from the start: the
  // collect launches below take its address when they are built
  {
    const int rc0 = g->ensure(0, K);
    if (rc0) return rc0;
  }
  // per handle: the arrays of its single-handle path (kkt_solver::bicgstab,
  // run_tri_solve)
  struct HV {
    kkt_solver* h;
    double *r, *rh, *p, *v, *s, *t, *y, *z, *x, *sc, *pa, *pb;
    int nbs, nbv;
    T *bp, *fp, *yy, *xp, *wv, *u;
    const T *E0, *E1;
    const S* PG;
    double bn;
    int it;
  };
  std::vector<HV> H(K);
  for (int q = 0; q < K; q++) {
    kkt_solver* h = hs[q];
    HV& a = H[q];
    a.h = h;
    a.r = h->vec(0); a.rh = h->vec(1); a.p = h->vec(2); a.v = h->vec(3); a.s = h->vec(4);
    a.t = h->vec(5); a.y = h->vec(6); a.z = h->vec(7); a.x = h->vec(8);
    a.sc = h->d_scal;
    a.nbs = spmv_blocks(h->n);
    a.nbv = vec_blocks(h->n);
    a.pa = h->d_part2;
    a.pb = h->d_part2 + std::max(a.nbs, 1);
    T* sv = (T*)h->d_svec;
    const size_t n = (size_t)h->n;
    a.bp = sv; a.fp = sv + n; a.yy = sv + 2 * n; a.xp = sv + 3 * n; a.wv = sv + 4 * n;
    T* F = (T*)(h->fprec == KKT_PREC_FP64 ? (void*)h->d_F64 : (void*)h->d_F32);
    T* PK = (T*)(h->fprec == KKT_PREC_FP64 ? (void*)h->d_PK64 : (void*)h->d_PK32);
    a.u = h->pk_slot(PK);
    a.E0 = F + h->front_elems;
    a.E1 = a.E0 + h->n;
    a.PG = (h->solve_q(true) == 2) ? (const S*)h->d_PK16 : (const S*)PK;
    a.bn = 1.0;
    a.it = 0;
    status[q] = 0;
  }
  // the triangular solve's core (forward levels, D^-1, backward levels) for a
  // set of handles: bitwise the launches of run_tri_solve, level by level
  auto core = [&](const std::vector<int>& act, std::vector<GLaunch>& L) {
    int maxl = 0;
    for (int q : act) maxl = std::max(maxl, H[q].h->S.nlevels);
    for (int Lv = 0; Lv < maxl; Lv++) {
      GLaunch g1, g2;
      g1.fam = g2.fam = GF_FWD;
      for (int q : act) {
        kkt_solver* h = H[q].h;
        if (Lv >= h->S.nlevels || !h->ff_l[Lv].cnt) continue;
        SegFwd f;
        f.d = h->ds; f.F = h->pk_fwd(H[q].PG); f.bp = H[q].bp; f.fp = H[q].fp; f.y = H[q].yy; f.u = H[q].u;
        f.items = h->d_ff + h->ff_l[Lv].off;
        g1.add(f, h->ff_l[Lv].cnt);
        if (h->fs_l[Lv].cnt) {
          f.items = h->d_ff + h->fs_l[Lv].off;
          g2.add(f, h->fs_l[Lv].cnt);
        }
      }
      L.push_back(std::move(g1));
      L.push_back(std::move(g2));
    }
    GLaunch gd;
    for (int q : act) {
      SegVec v = vop(SEG_DSOLVE, H[q].h->n);
      v.tin = H[q].E0; v.tin2 = H[q].E1; v.piv = H[q].h->d_piv; v.wpos = H[q].h->ds.wpos;
      v.tin3 = H[q].yy; v.tout = H[q].wv;
      gd.add(v, vblk(H[q].h->n));
    }
    L.push_back(std::move(gd));
    for (int j = 0; j < maxl; j++) {
      GLaunch gb;
      gb.fam = GF_BWD;
      for (int q : act) {
        kkt_solver* h = H[q].h;
        const int Lv = h->S.nlevels - 1 - j;
        if (Lv < 0 || !h->bc_l[Lv].cnt) continue;
        SegBwd b;
        b.d = h->ds; b.F = h->pk_bwd(H[q].PG); b.wv = H[q].wv; b.xp = H[q].xp;
        b.items = h->d_bc + h->bc_l[Lv].off;
        gb.add(b, h->bc_l[Lv].cnt);
      }
      L.push_back(std::move(gb));
    }
  };
  // uploads the core tables (kept while the active set is unchanged) and a
  // phase's own launches, runs pre + core + post, gathers the scalars of `act`
  // and waits for them
  std::vector<GLaunch> coreL;
  std::vector<int> core_set;
  size_t core_end = 0;  // 0: the core tables are not (or no longer) in the device blob
  int core_gen = -1;
  auto run_phase = [&](const std::vector<int>& act, std::vector<GLaunch>& pre, bool with_core,
                       std::vector<GLaunch>& post) -> int {
    if (with_core && act != core_set) {
      coreL.clear();
      core(act, coreL);
      core_set = act;
      core_end = 0;
    }
    if ((int)act.size() > g->sc_cap || !g->d_sc)
      return g->fail(KKT_ERR_ARGUMENT, "group scalar buffer smaller than the group");
    GLaunch col;
    for (size_t i = 0; i < act.size(); i++) {
      SegVec v = vop(SEG_COLLECT, 16);
      v.a = H[act[i]].sc;
      v.d = g->d_sc + 16 * i;
      col.add(v, 1);
    }
    post.push_back(std::move(col));
    if (!segs_fit(coreL) || !segs_fit(pre) || !segs_fit(post))
      return g->fail(KKT_ERR_ARGUMENT, "group launch exceeds SEG_MAX segments");
    const size_t need_core = with_core ? pack_launches(coreL, nullptr, 0) : 0;
    size_t need = pack_launches(pre, nullptr, need_core);
    need = pack_launches(post, nullptr, need);
    int rc = g->ensure(need, (int)act.size());
    if (rc) return rc;
    if (core_gen != g->gen) core_end = 0;  // a reallocated blob lost the core tables
    if (with_core && core_end == 0) {
      core_end = pack_launches(coreL, g->h_blob, 0);
      core_gen = g->gen;
      if (hipMemcpyAsync(g->d_blob, g->h_blob, core_end, hipMemcpyHostToDevice, st) != hipSuccess)
        return g->fail(KKT_ERR_COPY, "group table upload failed");
    }
    const size_t b0 = (need_core + 15) & ~(size_t)15;
    size_t e = pack_launches(pre, g->h_blob, b0);
    e = pack_launches(post, g->h_blob, e);
    if (e > b0 && hipMemcpyAsync(g->d_blob + b0, g->h_blob + b0, e - b0, hipMemcpyHostToDevice, st) != hipSuccess)
      return g->fail(KKT_ERR_COPY, "group table upload failed");
    run_launches<T, S>(pre, g->d_blob, st);
    if (with_core) run_launches<T, S>(coreL, g->d_blob, st);
    run_launches<T, S>(post, g->d_blob, st);
    if (hipMemcpyAsync(g->h_sc, g->d_sc, (size_t)16 * act.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return g->fail(KKT_ERR_REDUCTION, "group step failed");
    return KKT_OK;
  };
  auto finish = [&](int q, int code, const char* msg) {
    status[q] = code;
    if (code <= 0) H[q].h->fail(code, msg);
    else H[q].h->ok();
  };

  // ---- start: new values, x0 = M^-1 b, r = b - A x0, |b|, |r| ----
  std::vector<int> act(K);
  for (int q = 0; q < K; q++) act[q] = q;
  {
    std::vector<GLaunch> pre, post;
    GLaunch cp, ex, ga;
    for (int q : act) {
      kkt_solver* h = H[q].h;
      if (vin[q] != h->d_vals) {
        SegVec v = vop(SEG_COPY, h->nnz);
        v.a = vin[q]; v.d = h->d_vals;
        cp.add(v, vblk(h->nnz));
      }
      SegVec e = vop(SEG_EXPAND, h->nnzf);
      e.idx = h->d_map_f; e.a = h->d_vals; e.d = h->d_valsf;
      ex.add(e, vblk(h->nnzf));
      SegVec gg = vop(SEG_GATHER, h->n);
      gg.idx = h->d_perm; gg.a = bin[q]; gg.tout = H[q].bp;
      ga.add(gg, vblk(h->n));
    }
    pre.push_back(std::move(cp));
    pre.push_back(std::move(ex));
    pre.push_back(std::move(ga));
    GLaunch sc, sp, dp, rf;
    sp.fam = GF_SPMV;
    dp.fam = rf.fam = GF_RED;
    for (int q : act) {
      kkt_solver* h = H[q].h;
      SegVec v = vop(SEG_SCATTER, h->n);
      v.idx = h->d_perm; v.tin = H[q].xp; v.d = H[q].x;
      sc.add(v, vblk(h->n));
      SegSpmv m{};
      m.n = h->n; m.mode = 1; m.rowptr = h->d_rowptr_f; m.col = h->d_col_f; m.val = h->d_valsf;
      m.x = H[q].x; m.b = bin[q]; m.y = H[q].r;
      sp.add(m, spmv_blocks(h->n));
      // |b|^2 into slot 9 and |r|^2 into slot 8, each through its own half of
      // the handle's partials
      SegRed d1{}, d2{};
      d1.op = d2.op = SEG_DOTP;
      d1.n = d2.n = h->n;
      d1.p1 = d1.p2 = bin[q]; d1.out1 = h->d_partial;
      d2.p1 = d2.p2 = H[q].r; d2.out1 = h->d_partial + red_blocks();
      dp.add(d1, red_blocks());
      dp.add(d2, red_blocks());
      SegRed f1{}, f2{};
      f1.op = f2.op = SEG_REDF;
      f1.np = f2.np = red_blocks();
      f1.p1 = h->d_partial; f1.out1 = H[q].sc + 9;
      f2.p1 = h->d_partial + red_blocks(); f2.out1 = H[q].sc + 8;
      rf.add(f1, 1);
      rf.add(f2, 1);
    }
    post.push_back(std::move(sc));
    post.push_back(std::move(sp));
    post.push_back(std::move(dp));
    post.push_back(std::move(rf));
    int rc = run_phase(act, pre, true, post);
    if (rc) return rc;
    std::vector<int> next;
    GLaunch ini_rh, ini_p, ini_v, ini_sc;
    for (size_t i = 0; i < act.size(); i++) {
      const int q = act[i];
      const double* s = g->h_sc + 16 * i;
      double bn = std::sqrt(s[9]);
      const double rn = std::sqrt(s[8]);
      if (bn == 0.0) bn = 1.0;
      H[q].bn = bn;
      H[q].it = 1;
      if (rn / bn <= tolRel) {
        finish(q, 1, "");
        continue;
      }
      next.push_back(q);
      kkt_solver* h = H[q].h;
      SegVec c = vop(SEG_COPY, h->n);
      c.a = H[q].r; c.d = H[q].rh;
      ini_rh.add(c, vblk(h->n));
      SegVec f = vop(SEG_FILL, h->n);
      f.d = H[q].p;
      ini_p.add(f, vblk(h->n));
      f.d = H[q].v;
      ini_v.add(f, vblk(h->n));
      SegVec z = vop(SEG_BICG_INIT, 1);
      z.d = H[q].sc;
      ini_sc.add(z, 1);
    }
    act.swap(next);
    if (!act.empty()) {
      std::vector<GLaunch> L;
      L.push_back(std::move(ini_rh));
      L.push_back(std::move(ini_p));
      L.push_back(std::move(ini_v));
      L.push_back(std::move(ini_sc));
      const size_t b0 = (core_end + 15) & ~(size_t)15;
      size_t need = pack_launches(L, nullptr, b0);
      if ((rc = g->ensure(need, K))) return rc;
      const size_t e = pack_launches(L, g->h_blob, b0);
      if (hipMemcpyAsync(g->d_blob + b0, g->h_blob + b0, e - b0, hipMemcpyHostToDevice, st) != hipSuccess)
        return g->fail(KKT_ERR_COPY, "group table upload failed");
      run_launches<T, S>(L, g->d_blob, st);
      if (hipStreamSynchronize(st) != hipSuccess) return g->fail(KKT_ERR_REDUCTION, "group step failed");
    }
  }
  // ---- iterations ----
  for (int it = 1; it <= maxIter && !act.empty(); it++) {
    // head: p update + y = M^-1 p, v = A y, alpha, s = r - alpha v, |s|
    std::vector<GLaunch> pre, post;
    GLaunch gp, scat, sp, r1, bs, r2;
    sp.fam = GF_SPMV;
    r1.fam = r2.fam = GF_RED;
    for (int q : act) {
      kkt_solver* h = H[q].h;
      SegVec v = vop(SEG_GATHER_P, h->n);
      v.idx = h->d_perm; v.sc = H[q].sc; v.a = H[q].r; v.b = H[q].v; v.d = H[q].p; v.tout = H[q].bp;
      gp.add(v, vblk(h->n));
      SegVec s = vop(SEG_SCATTER, h->n);
      s.idx = h->d_perm; s.tin = H[q].xp; s.d = H[q].y;
      scat.add(s, vblk(h->n));
      SegSpmv m{};
      m.n = h->n; m.mode = 0; m.rowptr = h->d_rowptr_f; m.col = h->d_col_f; m.val = h->d_valsf;
      m.x = H[q].y; m.y = H[q].v; m.a1 = H[q].rh; m.part1 = H[q].pa;
      sp.add(m, H[q].nbs);
      SegRed a{};
      a.op = SEG_RED2; a.p1 = H[q].pa; a.np = H[q].nbs; a.out1 = H[q].sc + 4;
      r1.add(a, 1);
      SegVec b = vop(SEG_BICG_S, h->n);
      b.sc = H[q].sc; b.a = H[q].r; b.b = H[q].v; b.d = H[q].s; b.part1 = H[q].pa;
      bs.add(b, vblk(h->n));
      SegRed c{};
      c.op = SEG_RED2; c.p1 = H[q].pa; c.np = H[q].nbv; c.out1 = H[q].sc + 5;
      r2.add(c, 1);
    }
    pre.push_back(std::move(gp));
    post.push_back(std::move(scat));
    post.push_back(std::move(sp));
    post.push_back(std::move(r1));
    post.push_back(std::move(bs));
    post.push_back(std::move(r2));
    int rc = run_phase(act, pre, true, post);
    if (rc) return rc;
    std::vector<int> tail;
    GLaunch xa;
    for (size_t i = 0; i < act.size(); i++) {
      const int q = act[i];
      const double* s = g->h_sc + 16 * i;
      H[q].it = it;
      if (std::fabs(s[0]) < eps) { finish(q, KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: rho ~ 0"); continue; }
      if (std::fabs(s[4]) < eps) { finish(q, KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: (r_hat,v) ~ 0"); continue; }
      if (std::sqrt(s[5]) / H[q].bn <= tolRel) {
        SegVec v = vop(SEG_BICG_XA, H[q].h->n);
        v.sc = H[q].sc; v.a = H[q].y; v.d = H[q].x;
        xa.add(v, vblk(H[q].h->n));
        finish(q, it, "");
        continue;
      }
      tail.push_back(q);
    }
    if (xa.nblk()) {
      std::vector<GLaunch> L;
      L.push_back(std::move(xa));
      const size_t b0 = (core_end + 15) & ~(size_t)15;
      if ((rc = g->ensure(pack_launches(L, nullptr, b0), K))) return rc;
      const size_t e = pack_launches(L, g->h_blob, b0);
      if (hipMemcpyAsync(g->d_blob + b0, g->h_blob + b0, e - b0, hipMemcpyHostToDevice, st) != hipSuccess)
        return g->fail(KKT_ERR_COPY, "group table upload failed");
      run_launches<T, S>(L, g->d_blob, st);
      // the staging buffer is rewritten by the next phase only after this
      // upload has been read
      if (hipStreamSynchronize(st) != hipSuccess) return g->fail(KKT_ERR_REDUCTION, "group step failed");
    }
    act.swap(tail);
    if (act.empty()) break;
    // tail: z = M^-1 s, t = A z, omega, x and r updates, |r|, next rho
    pre.clear();
    post.clear();
    GLaunch ga, sc2, sp2, r3, xr, r4;
    sp2.fam = GF_SPMV;
    r3.fam = r4.fam = GF_RED;
    for (int q : act) {
      kkt_solver* h = H[q].h;
      SegVec v = vop(SEG_GATHER, h->n);
      v.idx = h->d_perm; v.a = H[q].s; v.tout = H[q].bp;
      ga.add(v, vblk(h->n));
      SegVec s = vop(SEG_SCATTER, h->n);
      s.idx = h->d_perm; s.tin = H[q].xp; s.d = H[q].z;
      sc2.add(s, vblk(h->n));
      SegSpmv m{};
      m.n = h->n; m.mode = 0; m.rowptr = h->d_rowptr_f; m.col = h->d_col_f; m.val = h->d_valsf;
      m.x = H[q].z; m.y = H[q].t; m.a1 = H[q].s; m.part1 = H[q].pa; m.part2 = H[q].pb;
      sp2.add(m, H[q].nbs);
      SegRed a{};
      a.op = SEG_RED2; a.p1 = H[q].pa; a.p2 = H[q].pb; a.np = H[q].nbs; a.out1 = H[q].sc + 6; a.out2 = H[q].sc + 7;
      r3.add(a, 1);
      SegVec b = vop(SEG_BICG_XR, h->n);
      b.sc = H[q].sc; b.a = H[q].y; b.b = H[q].z; b.c = H[q].s; b.e = H[q].t; b.d = H[q].x; b.f = H[q].r;
      b.g = H[q].rh; b.part1 = H[q].pa; b.part2 = H[q].pb;
      xr.add(b, vblk(h->n));
      SegRed c{};
      c.op = SEG_RED2; c.p1 = H[q].pa; c.p2 = H[q].pb; c.np = H[q].nbv; c.out1 = H[q].sc + 8; c.out2 = H[q].sc + 10;
      c.sc = H[q].sc;
      r4.add(c, 1);
    }
    pre.push_back(std::move(ga));
    post.push_back(std::move(sc2));
    post.push_back(std::move(sp2));
    post.push_back(std::move(r3));
    post.push_back(std::move(xr));
    post.push_back(std::move(r4));
    if ((rc = run_phase(act, pre, true, post))) return rc;
    std::vector<int> next;
    for (size_t i = 0; i < act.size(); i++) {
      const int q = act[i];
      const double* s = g->h_sc + 16 * i + 6;
      if (s[1] < eps) { finish(q, KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: (t,t) ~ 0"); continue; }
      if (std::sqrt(s[2]) / H[q].bn <= tolRel) { finish(q, it, ""); continue; }
      if (std::fabs(s[0] / s[1]) < eps) { finish(q, KKT_ERR_ITER_BREAKDOWN, "BiCGStab breakdown: omega ~ 0"); continue; }
      if (it == maxIter) {
        finish(q, KKT_ERR_ITER_NO_CONVERGE, "BiCGStab did not converge within iteration limit");
        continue;
      }
      next.push_back(q);
    }
    act.swap(next);
  }
  for (int q : act) finish(q, KKT_ERR_ITER_NO_CONVERGE, "BiCGStab did not converge within iteration limit");
  // converged handles: x to the caller's vector
  GLaunch out;
  for (int q = 0; q < K; q++) {
    if (status[q] <= 0) continue;
    SegVec v = vop(SEG_COPY, H[q].h->n);
    v.a = H[q].x; v.d = xout[q];
    out.add(v, vblk(H[q].h->n));
  }
  if (out.nblk()) {
    std::vector<GLaunch> L;
    L.push_back(std::move(out));
    int rc;
    if ((rc = g->ensure(pack_launches(L, nullptr, 0), K))) return rc;
    const size_t e = pack_launches(L, g->h_blob, 0);
    if (hipMemcpyAsync(g->d_blob, g->h_blob, e, hipMemcpyHostToDevice, st) != hipSuccess)
      return g->fail(KKT_ERR_COPY, "group table upload failed");
    run_launches<T, S>(L, g->d_blob, st);
  }
  if (hipStreamSynchronize(st) != hipSuccess) return g->fail(KKT_ERR_SOLVE, "group solve failed");
  return KKT_OK;
}

// =====================================================================
// Batched systems: many independent KKT systems of uneven size in one handle.
// The systems are concatenated into one block-diagonal upper CSR whose
// assembly forest holds one tree per system (one per connected component),
// so every factor and triangular-solve launch covers all systems at one tree
// level and each is one captured graph.  Pivot perturbation thresholds and
// inertia are per component; refinement and its stopping test are per system
// (segmented residual norms, corrections applied only to systems still
// refining), so a system's result does not depend on the rest of the batch.
// =====================================================================
struct kkt_batch {
  kkt_solver* h = nullptr;
  int nsys = 0;
  std::vector<int> sys_n;
  std::vector<long long> row_off, val_off;  // nsys + 1 each
  // per-system row chunks (at most 8192 rows) for the segmented norms
  long long *d_rca = nullptr, *d_rcb = nullptr;
  int *d_rcptr = nullptr, *d_row_sys = nullptr;
  double *d_rpart = nullptr, *d_rnorm = nullptr;
  unsigned char* d_active = nullptr;
  int nrchunks = 0;
  std::vector<int> refine_steps;
  std::vector<double> relres;
  std::string err;
  bool has_err = false;

  int fail(int code, const std::string& m) {
    err = m;
    has_err = true;
    return code;
  }
  int from_h(int rc) {
    if (rc) {
      const char* m = kkt_last_error(h);
      return fail(rc, m ? m : "batch operation failed");
    }
    has_err = false;
    return KKT_OK;
  }
  void free_dev() {
    DevOpLock lk(g_capture_mu);
    for (void* q : {(void*)d_rca, (void*)d_rcb, (void*)d_rcptr, (void*)d_row_sys, (void*)d_rpart,
                    (void*)d_rnorm, (void*)d_active})
      if (q) (void)hipFree(q);
    d_rca = d_rcb = nullptr;
    d_rcptr = d_row_sys = nullptr;
    d_rpart = d_rnorm = nullptr;
    d_active = nullptr;
  }
  ~kkt_batch() {
    if (h && h->dev_ok) (void)hipStreamSynchronize(h->st);
    free_dev();
    delete h;
  }

  int set_systems(int ns, const int* sysN, const int* ro, const int* ci, int mtype) {
    if (ns <= 0 || !sysN || !ro || !ci) return fail(KKT_ERR_ARGUMENT, "bad batch arrays");
    std::vector<int> gro(1, 0), gci;
    std::vector<int> sn(ns);
    std::vector<long long> roff(ns + 1, 0), voff(ns + 1, 0);
    long long rpos = 0, cpos = 0;
    for (int q = 0; q < ns; q++) {
      const int n = sysN[q];
      if (n <= 0) return fail(KKT_ERR_ARGUMENT, "system " + std::to_string(q) + " has no rows");
      const int* r = ro + rpos;
      const int nnz = r[n] - r[0];
      if (r[0] != 0) return fail(KKT_ERR_PATTERN, "system " + std::to_string(q) + ": rowOffs must start at 0");
      std::string e;
      if (validate_pattern(n, nnz, r, ci + cpos, e) != 0)
        return fail(KKT_ERR_PATTERN, "system " + std::to_string(q) + ": " + e);
      if (roff[q] + n > INT32_MAX || (long long)gro.back() + nnz > INT32_MAX)
        return fail(KKT_ERR_ARGUMENT, "batch larger than 2^31 rows or entries");
      const int base = (int)roff[q], vbase = gro.back();
      for (int i = 1; i <= n; i++) gro.push_back(vbase + r[i]);
      for (int p = 0; p < nnz; p++) gci.push_back(base + ci[cpos + p]);
      sn[q] = n;
      roff[q + 1] = roff[q] + n;
      voff[q + 1] = voff[q] + nnz;
      rpos += n + 1;
      cpos += nnz;
    }
    if (!h && !(h = kkt_create())) return fail(KKT_ERR_INIT, "handle creation failed");
    free_dev();
    int rc = from_h(kkt_set_pattern(h, (int)roff[ns], (int)gci.size(), gro.data(), gci.data(), mtype));
    if (rc) return rc;
    nsys = ns;
    sys_n = sn;
    row_off = roff;
    val_off = voff;
    refine_steps.assign(ns, 0);
    relres.assign(ns, 0.0);
    return KKT_OK;
  }

  int ensure_dev() {
    if (d_rca) return KKT_OK;
    std::vector<long long> ca, cb;
    std::vector<int> cptr(nsys + 1, 0), rsys(row_off[nsys]);
    for (int q = 0; q < nsys; q++) {
      for (long long a = row_off[q]; a < row_off[q + 1]; a += 8192) {
        ca.push_back(a);
        cb.push_back(std::min(row_off[q + 1], a + 8192));
      }
      cptr[q + 1] = (int)ca.size();
      for (long long i = row_off[q]; i < row_off[q + 1]; i++) rsys[i] = q;
    }
    nrchunks = (int)ca.size();
    DevOpLock lk(g_capture_mu);
    bool ok = hipMalloc(&d_rca, ca.size() * sizeof(long long)) == hipSuccess &&
              hipMalloc(&d_rcb, cb.size() * sizeof(long long)) == hipSuccess &&
              hipMalloc(&d_rcptr, cptr.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&d_row_sys, rsys.size() * sizeof(int)) == hipSuccess &&
              hipMalloc(&d_rpart, ca.size() * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_rnorm, nsys * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_active, nsys) == hipSuccess &&
              hipMemcpy(d_rca, ca.data(), ca.size() * sizeof(long long), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_rcb, cb.data(), cb.size() * sizeof(long long), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_rcptr, cptr.data(), cptr.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(d_row_sys, rsys.data(), rsys.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) return fail(KKT_ERR_ALLOC, "device allocation failed (batch norms)");
    return KKT_OK;
  }

  // squared 2-norms of each system's slice of v (one synchronization)
  int seg_norms(const double* v, std::vector<double>& out) {
    launch_seg_sumsq(nrchunks, d_rca, d_rcb, v, d_rpart, nsys, d_rcptr, d_rnorm, h->st);
    out.resize(nsys);
    if (hipMemcpyAsync(out.data(), d_rnorm, nsys * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess)
      return fail(KKT_ERR_REDUCTION, "segmented reduction failed");
    return KKT_OK;
  }

  // Refined solve of every system, with the single-system rules per system:
  // FP32 factors (and the FP64 factor of a fallback) refine each system to
  // refine_tol; a stall in any system refactors the batch in FP64 (once) and
  // the refinement restarts from the FP64 solve; a stall on the FP64 factor
  // returns KKT_ERR_REFINE_STALL.  Plain FP64 factors take ir_steps steps.
  int solve_device(const double* d_b, double* d_x) {
    kkt_solver& S = *h;
    if (S.state != kkt_solver::FACTORED) return fail(KKT_ERR_STATE, "solve called before factor");
    int rc = ensure_dev();
    if (rc) return rc;
    S.tri_solve(d_b, d_x);
    std::fill(refine_steps.begin(), refine_steps.end(), 0);
    std::fill(relres.begin(), relres.end(), 0.0);
    const bool to_tol = S.fprec == KKT_PREC_FP32_REFINE || S.fallback_active;
    const int want = to_tol ? S.refine_max : S.ir_steps;
    if (want > 0) {
      double* r = S.vec(0);
      double* dx = S.vec(1);
      std::vector<double> bn, rn, prev(nsys, 0.0);
      std::vector<int> stall(nsys, 0);
      std::vector<unsigned char> active(nsys, 1);
      if ((rc = seg_norms(d_b, bn))) return rc;
      for (double& v : bn) v = (v == 0.0) ? 1.0 : std::sqrt(v);
      for (int step = 0;; step++) {
        S.spmv(d_x, d_b, r, 1);
        if ((rc = seg_norms(r, rn))) return rc;
        int nact = 0, stalled = -1;
        for (int q = 0; q < nsys; q++) {
          if (!active[q]) continue;
          const double rel = std::sqrt(rn[q]) / bn[q];
          relres[q] = rel;
          if (to_tol) {
            if (rel <= S.refine_tol) { active[q] = 0; continue; }
            if (step > 0 && rel > 0.5 * prev[q]) stall[q]++; else stall[q] = 0;
            if (step >= S.refine_max || stall[q] >= 2 || !std::isfinite(rel)) {
              if (stalled < 0) stalled = q;
            }
          } else if (step >= S.ir_steps) {
            active[q] = 0;
            continue;
          }
          prev[q] = rel;
          nact++;
        }
        if (stalled >= 0) {
          if (S.fprec == KKT_PREC_FP32_REFINE) {
            if ((rc = from_h(S.fallback_to_fp64() ? S.fail(KKT_ERR_FACTORIZATION, S.err) : KKT_OK))) return rc;
            S.tri_solve(d_b, d_x);
            std::fill(active.begin(), active.end(), 1);
            std::fill(stall.begin(), stall.end(), 0);
            std::fill(prev.begin(), prev.end(), 0.0);
            std::fill(refine_steps.begin(), refine_steps.end(), 0);
            step = -1;
            continue;
          }
          (void)hipStreamSynchronize(S.st);
          char msg[200];
          snprintf(msg, sizeof msg, "refinement of system %d stalled at relative residual %.3g > %.3g after the "
                   "FP64 refactor", stalled, relres[stalled], S.refine_tol);
          return fail(KKT_ERR_REFINE_STALL, msg);
        }
        if (nact == 0) break;
        if (hipMemcpyAsync(d_active, active.data(), nsys, hipMemcpyHostToDevice, S.st) != hipSuccess)
          return fail(KKT_ERR_COPY, "hipMemcpyAsync (active systems) failed");
        S.tri_solve(r, dx);
        launch_axpy_masked(row_off[nsys], d_row_sys, d_active, dx, d_x, S.st);
        for (int q = 0; q < nsys; q++) refine_steps[q] += active[q];
      }
    }
    if (hipStreamSynchronize(S.st) != hipSuccess) return fail(KKT_ERR_SOLVE, "solve kernels failed");
    has_err = false;
    return KKT_OK;
  }

  int system_info(int q, kkt_batch_sysinfo* o) {
    if (q < 0 || q >= nsys || !o) return fail(KKT_ERR_ARGUMENT, "system index out of range");
    memset(o, 0, sizeof(*o));
    o->n = sys_n[q];
    o->refine_steps = refine_steps[q];
    o->last_rel_residual = relres[q];
    const kkt_solver& S = *h;
    if (S.state < kkt_solver::ANALYZED) return fail(KKT_ERR_STATE, "no analysis");
    if (S.state == kkt_solver::FACTORED && (int)S.h_rctr.size() == 4 * S.nroots) {
      std::vector<char> seen(S.nroots, 0);
      for (long long i = row_off[q]; i < row_off[q + 1]; i++) {
        const int r = S.root_of_row[i];
        if (seen[r]) continue;
        seen[r] = 1;
        o->npos += S.h_rctr[4 * r];
        o->nneg += S.h_rctr[4 * r + 1];
        o->nperturbed += S.h_rctr[4 * r + 2];
        o->n2x2 += S.h_rctr[4 * r + 3];
      }
    }
    has_err = false;
    return KKT_OK;
  }
};

// =====================================================================
// C ABI
// =====================================================================
template <typename T>
static int cmp_dev(const T* d, const std::vector<T>& hv) {
  std::vector<T> tmp(hv.size());
  DevOpLock lk(g_capture_mu);
  if (!hv.empty() && hipMemcpy(tmp.data(), d, hv.size() * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) return 1;
  return memcmp(tmp.data(), hv.data(), hv.size() * sizeof(T)) != 0;
}

#define CHECK_H(h) \
  if (!(h)) return KKT_ERR_ARGUMENT

extern "C" {

kkt_solver* kkt_create(void) {
  try {
    return new kkt_solver();
  } catch (...) {
    return nullptr;
  }
}

void kkt_destroy(kkt_solver* h) { delete h; }

int kkt_set_pattern(kkt_solver* h, int n, int nnz, const int* rowOffs, const int* colIdxs, int mtype) {
  CHECK_H(h);
  int rc = h->set_pattern(n, nnz, rowOffs, colIdxs, mtype);
  if (!rc) h->ok();
  return rc;
}

int kkt_analyze(kkt_solver* h) {
  CHECK_H(h);
  int rc = h->analyze(nullptr);
  if (!rc) h->ok();
  return rc;
}

int kkt_analyze_values(kkt_solver* h, const double* vals) {
  CHECK_H(h);
  if (!vals) return h->fail(KKT_ERR_ARGUMENT, "null values");
  int rc = h->analyze(vals);
  if (!rc) h->ok();
  return rc;
}

int kkt_factor(kkt_solver* h, const double* vals) {
  CHECK_H(h);
  if (!vals) return h->fail(KKT_ERR_ARGUMENT, "null values");
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "factor called before analyze");
  if (hipMemcpyAsync(h->d_vals, vals, (size_t)h->nnz * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (vals H2D) failed");
  int rc = h->factor_device(h->d_vals, true);
  if (!rc) h->ok();
  return rc;
}

int kkt_factor_device(kkt_solver* h, const double* d_vals) {
  CHECK_H(h);
  if (!d_vals) return h->fail(KKT_ERR_ARGUMENT, "null values");
  int rc = h->factor_device(d_vals, false);
  if (!rc) h->ok();
  return rc;
}

int kkt_clear_device_values(kkt_solver* h) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  launch_fill(h->nnz, 0.0, h->d_vals, h->st);
  if (hipStreamSynchronize(h->st) != hipSuccess) return h->fail(KKT_ERR_DEVICE, "clear failed");
  h->ok();
  return KKT_OK;
}

int kkt_get_device_values(kkt_solver* h, double* vals) {
  CHECK_H(h);
  if (!vals) return h->fail(KKT_ERR_ARGUMENT, "null values");
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  if (hipMemcpyAsync(vals, h->d_vals, (size_t)h->nnz * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (device values) failed");
  h->ok();
  return KKT_OK;
}

int kkt_add_device_values(kkt_solver* h, const int* slots, const double* vals, int nvals, double scale) {
  CHECK_H(h);
  if (nvals < 0 || (nvals > 0 && (!slots || !vals))) return h->fail(KKT_ERR_ARGUMENT, "bad slot/value arrays");
  std::vector<const double*> dp;
  int rc = h->device_add(slots, nvals, scale, [&](double* c) -> int {
    if (nvals > 0 && hipMemcpyAsync(c, vals, (size_t)nvals * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
      return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (values) failed");
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_scaled_diagonal3_device_values(kkt_solver* h, const int* diagSlots, const double* masses,
                                           int nblocks, double scale) {
  CHECK_H(h);
  if (nblocks < 0 || (nblocks > 0 && (!diagSlots || !masses))) return h->fail(KKT_ERR_ARGUMENT, "bad block arrays");
  std::vector<const double*> dp;
  int rc = h->device_add(diagSlots, 3LL * nblocks, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{masses, (size_t)nblocks * sizeof(double)}}, dp);
    if (r) return r;
    launch_contrib_diag3(nblocks, dp[0], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_scaled_block3_device_values(kkt_solver* h, const int* blockSlots, const double* blockVals,
                                        const double* blockScales, int nblocks, double scale) {
  CHECK_H(h);
  if (nblocks < 0 || (nblocks > 0 && (!blockSlots || !blockVals || !blockScales)))
    return h->fail(KKT_ERR_ARGUMENT, "bad block arrays");
  std::vector<const double*> dp;
  int rc = h->device_add(blockSlots, 9LL * nblocks, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{blockVals, (size_t)nblocks * 9 * sizeof(double)},
                             {blockScales, (size_t)nblocks * sizeof(double)}}, dp);
    if (r) return r;
    launch_contrib_block3(nblocks, dp[0], dp[1], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_material_stiffness3_device_values(kkt_solver* h, const int* blockSlots, const double* gis,
                                              const double* gjs, const double* Ds, const double* sigmas,
                                              const double* dvs, int nblocks, double scale) {
  CHECK_H(h);
  if (nblocks < 0 || (nblocks > 0 && (!blockSlots || !gis || !gjs || !Ds || !sigmas || !dvs)))
    return h->fail(KKT_ERR_ARGUMENT, "bad descriptor arrays");
  std::vector<const double*> dp;
  int rc = h->device_add(blockSlots, 9LL * nblocks, scale, [&](double* c) -> int {
    const size_t nb = (size_t)nblocks * sizeof(double);
    int r = h->stage_inputs({{gis, 3 * nb}, {gjs, 3 * nb}, {Ds, 36 * nb}, {sigmas, 6 * nb}, {dvs, nb}}, dp);
    if (r) return r;
    launch_contrib_matstiff3(nblocks, dp[0], dp[1], dp[2], dp[3], dp[4], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

// ---- element-batched forms (CuDssJNI.cc / cudssBridgeKernels.cu:336-739) ----
// Shared checks of the compact element descriptors: offsets start at 0 and do not
// decrease, pair node indices lie inside their element, node target dimensions
// are 1..6, and each element's integration-point / gradient ranges fit.  (C++
// helpers: outside the C-linkage block.)
}  // extern "C"
namespace {
bool offsets_ok(const int* off, int nelems) {
  if (off[0] != 0) return false;
  for (int e = 0; e < nelems; e++)
    if (off[e + 1] < off[e]) return false;
  return true;
}
bool pairs_ok(const int* nodeCnt, const int* pairOff, const int* pairNodes, int nelems) {
  for (int e = 0; e < nelems; e++)
    for (int p = pairOff[e]; p < pairOff[e + 1]; p++) {
      const int i = pairNodes[2 * p], j = pairNodes[2 * p + 1];
      if (i < 0 || j < 0 || i >= nodeCnt[e] || j >= nodeCnt[e]) return false;
    }
  return true;
}
bool nodes_ok(const int* nodeCnt, const int* nodeOff, const int* nodeDims, int nelems) {
  for (int e = 0; e < nelems; e++)
    if (nodeCnt[e] < 0 || nodeOff[e + 1] - nodeOff[e] < nodeCnt[e]) return false;
  for (int q = 0; q < nodeOff[nelems]; q++)
    if (nodeDims[q] < 1 || nodeDims[q] > 6) return false;
  return true;
}
// the 36-slot block list with the entries past rowdim x coldim of every pair
// masked (they are not part of the target block)
std::vector<int> masked_slots36(const int* nodeOff, const int* pairOff, const int* pairNodes,
                                const int* blockSlots, const int* nodeDims, int nelems) {
  std::vector<int> m((size_t)36 * pairOff[nelems], -1);
  for (int e = 0; e < nelems; e++)
    for (int p = pairOff[e]; p < pairOff[e + 1]; p++) {
      const int n = nodeDims[nodeOff[e] + pairNodes[2 * p]] * nodeDims[nodeOff[e] + pairNodes[2 * p + 1]];
      for (int t = 0; t < n; t++) m[36 * (size_t)p + t] = blockSlots[36 * (size_t)p + t];
    }
  return m;
}
}  // namespace
extern "C" {

int kkt_add_material_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPairOffsets,
    const int* elemIpOffsets, const int* elemGradOffsets, const int* pairNodeIdxs, const int* blockSlots,
    const int* nodeDims, const double* nodeTransforms, const double* grads, const double* Ds,
    const double* sigmas, const double* dvs, int nelems, double scale) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  if (nelems < 0 || (nelems > 0 && (!elemNodeCounts || !elemNodeOffsets || !elemPairOffsets ||
                                    !elemIpOffsets || !elemGradOffsets)))
    return h->fail(KKT_ERR_ARGUMENT, "bad element arrays");
  if (nelems == 0) { h->ok(); return KKT_OK; }
  if (!offsets_ok(elemNodeOffsets, nelems) || !offsets_ok(elemPairOffsets, nelems) ||
      !offsets_ok(elemIpOffsets, nelems) || !offsets_ok(elemGradOffsets, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact material stiffness descriptor arrays are inconsistent");
  const long long nn = elemNodeOffsets[nelems], np = elemPairOffsets[nelems], ni = elemIpOffsets[nelems],
                  ng = elemGradOffsets[nelems];
  if ((np > 0 && (!pairNodeIdxs || !blockSlots)) || (nn > 0 && (!nodeDims || !nodeTransforms)) ||
      (ni > 0 && (!Ds || !sigmas || !dvs)) || (ng > 0 && !grads))
    return h->fail(KKT_ERR_ARGUMENT, "bad descriptor arrays");
  if (!nodes_ok(elemNodeCounts, elemNodeOffsets, nodeDims, nelems) ||
      !pairs_ok(elemNodeCounts, elemPairOffsets, pairNodeIdxs, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact material stiffness descriptor arrays are inconsistent");
  for (int e = 0; e < nelems; e++)
    if (elemGradOffsets[e + 1] - elemGradOffsets[e] <
        (long long)(elemIpOffsets[e + 1] - elemIpOffsets[e]) * elemNodeCounts[e])
      return h->fail(KKT_ERR_ARGUMENT, "fewer gradients than integration points x nodes");
  const std::vector<int> slots =
      masked_slots36(elemNodeOffsets, elemPairOffsets, pairNodeIdxs, blockSlots, nodeDims, nelems);
  std::vector<const double*> dp;
  const size_t I = sizeof(int), D = sizeof(double), ne = (size_t)nelems;
  int rc = h->device_add(slots.data(), 36 * np, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{elemNodeCounts, ne * I}, {elemNodeOffsets, (ne + 1) * I},
                             {elemPairOffsets, (ne + 1) * I}, {elemIpOffsets, (ne + 1) * I},
                             {elemGradOffsets, (ne + 1) * I}, {pairNodeIdxs, 2 * np * I},
                             {nodeDims, nn * I}, {nodeTransforms, 18 * nn * D}, {grads, 3 * ng * D},
                             {Ds, 36 * ni * D}, {sigmas, 6 * ni * D}, {dvs, ni * D}}, dp);
    if (r) return r;
    launch_contrib_matstiff3_elem(nelems, (const int*)dp[0], (const int*)dp[1], (const int*)dp[2],
                                  (const int*)dp[3], (const int*)dp[4], (const int*)dp[5],
                                  (const int*)dp[6], dp[7], dp[8], dp[9], dp[10], dp[11], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_linear_elastic_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemPairOffsets, const int* elemIpOffsets,
    const int* elemGradOffsets, const int* pairNodeIdxs, const int* blockSlots, const double* elemParams,
    const double* grads, const double* dvs, int nelems, double scale) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  if (nelems < 0 || (nelems > 0 && (!elemNodeCounts || !elemPairOffsets || !elemIpOffsets ||
                                    !elemGradOffsets || !elemParams)))
    return h->fail(KKT_ERR_ARGUMENT, "bad element arrays");
  if (nelems == 0) { h->ok(); return KKT_OK; }
  if (!offsets_ok(elemPairOffsets, nelems) || !offsets_ok(elemIpOffsets, nelems) ||
      !offsets_ok(elemGradOffsets, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact linear elastic descriptor arrays are inconsistent");
  const long long np = elemPairOffsets[nelems], ni = elemIpOffsets[nelems], ng = elemGradOffsets[nelems];
  if ((np > 0 && (!pairNodeIdxs || !blockSlots)) || (ni > 0 && !dvs) || (ng > 0 && !grads))
    return h->fail(KKT_ERR_ARGUMENT, "bad descriptor arrays");
  for (int e = 0; e < nelems; e++)
    if (elemNodeCounts[e] < 0 || elemGradOffsets[e + 1] - elemGradOffsets[e] <
                                     (long long)(elemIpOffsets[e + 1] - elemIpOffsets[e]) * elemNodeCounts[e])
      return h->fail(KKT_ERR_ARGUMENT, "fewer gradients than integration points x nodes");
  if (!pairs_ok(elemNodeCounts, elemPairOffsets, pairNodeIdxs, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact linear elastic descriptor arrays are inconsistent");
  std::vector<const double*> dp;
  const size_t I = sizeof(int), D = sizeof(double), ne = (size_t)nelems;
  int rc = h->device_add(blockSlots, 9 * np, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{elemNodeCounts, ne * I}, {elemPairOffsets, (ne + 1) * I},
                             {elemIpOffsets, (ne + 1) * I}, {elemGradOffsets, (ne + 1) * I},
                             {pairNodeIdxs, 2 * np * I}, {elemParams, 2 * ne * D}, {grads, 3 * ng * D},
                             {dvs, ni * D}}, dp);
    if (r) return r;
    launch_contrib_linelastic3_elem(nelems, (const int*)dp[0], (const int*)dp[1], (const int*)dp[2],
                                    (const int*)dp[3], (const int*)dp[4], dp[5], dp[6], dp[7], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_linear_elastic_stiffness3_element_geometry_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPairOffsets,
    const int* elemIpOffsets, const int* elemNaturalGradOffsets, const int* pairNodeIdxs,
    const int* blockSlots, const int* nodeDims, const double* nodeTransforms, const double* elemParams,
    const double* elemNodePositions, const double* naturalGrads, const double* ipWeights, int nelems,
    double scale) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  if (nelems < 0 || (nelems > 0 && (!elemNodeCounts || !elemNodeOffsets || !elemPairOffsets ||
                                    !elemIpOffsets || !elemNaturalGradOffsets || !elemParams)))
    return h->fail(KKT_ERR_ARGUMENT, "bad element arrays");
  if (nelems == 0) { h->ok(); return KKT_OK; }
  if (!offsets_ok(elemNodeOffsets, nelems) || !offsets_ok(elemPairOffsets, nelems) ||
      !offsets_ok(elemIpOffsets, nelems) || !offsets_ok(elemNaturalGradOffsets, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact linear elastic geometry descriptor arrays are inconsistent");
  const long long nn = elemNodeOffsets[nelems], np = elemPairOffsets[nelems], ni = elemIpOffsets[nelems],
                  ng = elemNaturalGradOffsets[nelems];
  if ((np > 0 && (!pairNodeIdxs || !blockSlots)) ||
      (nn > 0 && (!nodeDims || !nodeTransforms || !elemNodePositions)) || (ni > 0 && !ipWeights) ||
      (ng > 0 && !naturalGrads))
    return h->fail(KKT_ERR_ARGUMENT, "bad descriptor arrays");
  if (!nodes_ok(elemNodeCounts, elemNodeOffsets, nodeDims, nelems) ||
      !pairs_ok(elemNodeCounts, elemPairOffsets, pairNodeIdxs, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact linear elastic geometry descriptor arrays are inconsistent");
  for (int e = 0; e < nelems; e++)
    if (elemNaturalGradOffsets[e + 1] - elemNaturalGradOffsets[e] <
        (long long)(elemIpOffsets[e + 1] - elemIpOffsets[e]) * elemNodeCounts[e])
      return h->fail(KKT_ERR_ARGUMENT, "fewer natural gradients than integration points x nodes");
  const std::vector<int> slots =
      masked_slots36(elemNodeOffsets, elemPairOffsets, pairNodeIdxs, blockSlots, nodeDims, nelems);
  std::vector<const double*> dp;
  const size_t I = sizeof(int), D = sizeof(double), ne = (size_t)nelems;
  int rc = h->device_add(slots.data(), 36 * np, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{elemNodeCounts, ne * I}, {elemNodeOffsets, (ne + 1) * I},
                             {elemPairOffsets, (ne + 1) * I}, {elemIpOffsets, (ne + 1) * I},
                             {elemNaturalGradOffsets, (ne + 1) * I}, {pairNodeIdxs, 2 * np * I},
                             {nodeDims, nn * I}, {nodeTransforms, 18 * nn * D}, {elemParams, 2 * ne * D},
                             {elemNodePositions, 3 * nn * D}, {naturalGrads, 3 * ng * D},
                             {ipWeights, ni * D}}, dp);
    if (r) return r;
    launch_contrib_linelastic3_geom_elem(nelems, (const int*)dp[0], (const int*)dp[1], (const int*)dp[2],
                                         (const int*)dp[3], (const int*)dp[4], (const int*)dp[5],
                                         (const int*)dp[6], dp[7], dp[8], dp[9], dp[10], dp[11], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_add_dilational_stiffness3_element_device_values(
    kkt_solver* h, const int* elemNodeCounts, const int* elemNodeOffsets, const int* elemPressureCounts,
    const int* elemPairOffsets, const int* elemConstraintOffsets, const int* elemRinvOffsets,
    const int* pairNodeIdxs, const int* blockSlots, const int* nodeDims, const double* nodeTransforms,
    const double* constraints, const double* rinvs, int nelems, double scale) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  if (nelems < 0 || (nelems > 0 && (!elemNodeCounts || !elemNodeOffsets || !elemPressureCounts ||
                                    !elemPairOffsets || !elemConstraintOffsets || !elemRinvOffsets)))
    return h->fail(KKT_ERR_ARGUMENT, "bad element arrays");
  if (nelems == 0) { h->ok(); return KKT_OK; }
  if (!offsets_ok(elemNodeOffsets, nelems) || !offsets_ok(elemPairOffsets, nelems) ||
      !offsets_ok(elemConstraintOffsets, nelems) || !offsets_ok(elemRinvOffsets, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact dilational descriptor arrays are inconsistent");
  const long long nn = elemNodeOffsets[nelems], np = elemPairOffsets[nelems],
                  nc = elemConstraintOffsets[nelems], nr = elemRinvOffsets[nelems];
  if ((np > 0 && (!pairNodeIdxs || !blockSlots)) || (nn > 0 && (!nodeDims || !nodeTransforms)) ||
      (nc > 0 && !constraints) || (nr > 0 && !rinvs))
    return h->fail(KKT_ERR_ARGUMENT, "bad descriptor arrays");
  if (!nodes_ok(elemNodeCounts, elemNodeOffsets, nodeDims, nelems) ||
      !pairs_ok(elemNodeCounts, elemPairOffsets, pairNodeIdxs, nelems))
    return h->fail(KKT_ERR_ARGUMENT, "compact dilational descriptor arrays are inconsistent");
  for (int e = 0; e < nelems; e++) {
    const long long p = elemPressureCounts[e];
    if (p > 0 && (elemConstraintOffsets[e + 1] - elemConstraintOffsets[e] < p * elemNodeCounts[e] ||
                  elemRinvOffsets[e + 1] - elemRinvOffsets[e] < p * p))
      return h->fail(KKT_ERR_ARGUMENT, "constraint or Rinv range shorter than the pressure count needs");
  }
  const std::vector<int> slots =
      masked_slots36(elemNodeOffsets, elemPairOffsets, pairNodeIdxs, blockSlots, nodeDims, nelems);
  std::vector<const double*> dp;
  const size_t I = sizeof(int), D = sizeof(double), ne = (size_t)nelems;
  int rc = h->device_add(slots.data(), 36 * np, scale, [&](double* c) -> int {
    int r = h->stage_inputs({{elemNodeOffsets, (ne + 1) * I}, {elemPressureCounts, ne * I},
                             {elemPairOffsets, (ne + 1) * I}, {elemConstraintOffsets, (ne + 1) * I},
                             {elemRinvOffsets, (ne + 1) * I}, {pairNodeIdxs, 2 * np * I},
                             {nodeDims, nn * I}, {nodeTransforms, 18 * nn * D}, {constraints, 3 * nc * D},
                             {rinvs, nr * D}}, dp);
    if (r) return r;
    launch_contrib_dilational3_elem(nelems, (const int*)dp[0], (const int*)dp[1], (const int*)dp[2],
                                    (const int*)dp[3], (const int*)dp[4], (const int*)dp[5],
                                    (const int*)dp[6], dp[7], dp[8], dp[9], c, h->st);
    return KKT_OK;
  });
  if (!rc) h->ok();
  return rc;
}

int kkt_factor_device_values(kkt_solver* h) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "analyze() not previously called");
  int rc = h->factor_device(h->d_vals, true);
  if (!rc) h->ok();
  return rc;
}

int kkt_solve(kkt_solver* h, const double* b, double* x) {
  CHECK_H(h);
  if (!b || !x) return h->fail(KKT_ERR_ARGUMENT, "null vector");
  if (h->state != kkt_solver::FACTORED) return h->fail(KKT_ERR_STATE, "solve called before factor");
  double* db = h->vec(10);
  double* dx = h->vec(11);
  if (hipMemcpyAsync(db, b, (size_t)h->n * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (b) failed");
  int rc = h->solve_device(db, dx);
  if (rc) return rc;
  if (hipMemcpyAsync(x, dx, (size_t)h->n * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (x) failed");
  h->ok();
  return KKT_OK;
}

int kkt_solve_device(kkt_solver* h, const double* d_b, double* d_x) {
  CHECK_H(h);
  if (!d_b || !d_x || d_b == d_x) return h->fail(KKT_ERR_ARGUMENT, "null or aliased vectors");
  int rc = h->solve_device(d_b, d_x);
  if (!rc) h->ok();
  return rc;
}

int kkt_solve_device_noalloc(kkt_solver* h, const double* d_b, double* d_x) {
  CHECK_H(h);
  if (!d_b || !d_x || d_b == d_x) return h->fail(KKT_ERR_ARGUMENT, "null or aliased vectors");
  if (h->state != kkt_solver::FACTORED) return h->fail(KKT_ERR_STATE, "solve called before factor");
  unsigned char* scr = nullptr;
  hipError_t e = scratch_on_device(h->dev_id, &scr);
  if (e != hipSuccess) return h->fail(KKT_ERR_DEVICE, hip_msg("hipGetSymbolAddress (scratch) failed", e));
  int rc = h->solve_device_noalloc(d_b, d_x, scr, NOALLOC_SCRATCH_BYTES);
  if (!rc) h->ok();
  return rc;
}

int kkt_solve_multi(kkt_solver* h, int nrhs, const double* B, double* X) {
  CHECK_H(h);
  if (nrhs <= 0 || !B || !X) return h->fail(KKT_ERR_ARGUMENT, "nrhs must be positive");
  if (h->state != kkt_solver::FACTORED) return h->fail(KKT_ERR_STATE, "solve called before factor");
  int rc = h->ensure_multi();
  if (rc) return rc;
  double* dB = h->m_blk(2);
  double* dX = h->m_blk(3);
  for (int c0 = 0; c0 < nrhs; c0 += kkt_solver::MULTI_BLOCK) {
    const int nr = std::min(kkt_solver::MULTI_BLOCK, nrhs - c0);
    const size_t bytes = (size_t)nr * h->n * sizeof(double);
    if (hipMemcpyAsync(dB, B + (size_t)c0 * h->n, bytes, hipMemcpyHostToDevice, h->st) != hipSuccess)
      return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (B) failed");
    if ((rc = h->solve_multi_device(nr, dB, dX))) return rc;
    if (hipMemcpyAsync(X + (size_t)c0 * h->n, dX, bytes, hipMemcpyDeviceToHost, h->st) != hipSuccess ||
        hipStreamSynchronize(h->st) != hipSuccess)
      return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (X) failed");
  }
  h->ok();
  return KKT_OK;
}

int kkt_solve_multi_device(kkt_solver* h, int nrhs, const double* d_B, double* d_X) {
  CHECK_H(h);
  if (nrhs <= 0 || !d_B || !d_X) return h->fail(KKT_ERR_ARGUMENT, "nrhs must be positive");
  const size_t bytes = (size_t)nrhs * h->n * sizeof(double);
  const char *b0 = (const char*)d_B, *x0 = (const char*)d_X;
  if (b0 < x0 + bytes && x0 < b0 + bytes) return h->fail(KKT_ERR_ARGUMENT, "aliased right-hand sides and solutions");
  int rc = h->solve_multi_device(nrhs, d_B, d_X);
  if (!rc) h->ok();
  return rc;
}

int kkt_multiply(kkt_solver* h, const double* x, double* y) {
  CHECK_H(h);
  if (!x || !y) return h->fail(KKT_ERR_ARGUMENT, "null vector");
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "multiply called before analyze");
  double* dx = h->vec(10);
  double* dy = h->vec(11);
  if (hipMemcpyAsync(dx, x, (size_t)h->n * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (x) failed");
  // A is the matrix in the device value buffer (factor values or device assembly)
  int rc = h->set_values_device(h->d_vals);
  if (rc) return rc;
  h->spmv(dx, nullptr, dy, 0);
  if (hipMemcpyAsync(y, dy, (size_t)h->n * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return h->fail(KKT_ERR_SPMV, "multiply failed");
  h->tflush();
  h->ok();
  return KKT_OK;
}

int kkt_iterative_solve_device(kkt_solver* h, const double* d_vals, const double* d_b, double* d_x,
                               double tolRel, int maxIter) {
  CHECK_H(h);
  if (!d_vals || !d_b || !d_x) return h->fail(KKT_ERR_ARGUMENT, "null argument");
  if (h->state != kkt_solver::FACTORED) return h->fail(KKT_ERR_STATE, "iterativeSolve called before factor");
  if (maxIter <= 0) return h->fail(KKT_ERR_ARGUMENT, "maxIter must be positive");
  int rc = h->set_values_device(d_vals);
  if (rc) return rc;
  int it = 0;
  rc = h->bicgstab(d_b, d_x, tolRel, maxIter, &it);
  if (rc) return rc;
  h->ok();
  return it;
}

kkt_group* kkt_group_create(void) {
  kkt_group* g = new (std::nothrow) kkt_group();
  if (!g) return nullptr;
  if (hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking) != hipSuccess) {
    g->st = nullptr;
    delete g;
    return nullptr;
  }
  return g;
}

void kkt_group_destroy(kkt_group* g) { delete g; }

const char* kkt_group_last_error(kkt_group* g) {
  if (!g || !g->has_err) return nullptr;
  return g->err.c_str();
}

int kkt_group_iterative_solve_device(kkt_group* g, int nh, kkt_solver* const* hs, const double* const* d_vals,
                                     const double* const* d_b, double* const* d_x, double tolRel, int maxIter,
                                     int* status) {
  if (!g) return KKT_ERR_ARGUMENT;
  g->has_err = false;
  if (nh < 1 || nh > KKT_GROUP_MAX) return g->fail(KKT_ERR_ARGUMENT, "group size must be 1..KKT_GROUP_MAX");
  if (!hs || !d_vals || !d_b || !d_x || !status) return g->fail(KKT_ERR_ARGUMENT, "null argument");
  if (maxIter <= 0) return g->fail(KKT_ERR_ARGUMENT, "maxIter must be positive");
  for (int i = 0; i < nh; i++) {
    kkt_solver* h = hs[i];
    if (!h || !d_vals[i] || !d_b[i] || !d_x[i]) return g->fail(KKT_ERR_ARGUMENT, "null handle or array");
    for (int j = 0; j < i; j++)
      if (hs[j] == h) return g->fail(KKT_ERR_ARGUMENT, "a handle appears twice in the group");
    if (h->state != kkt_solver::FACTORED)
      return g->fail(KKT_ERR_STATE, "iterativeSolve called before factor (handle " + std::to_string(i) + ")");
  }
  // the handles' own streams are idle before the group writes their buffers
  for (int i = 0; i < nh; i++)
    if (hipStreamSynchronize(hs[i]->st) != hipSuccess) return g->fail(KKT_ERR_DEVICE, "handle stream failed");
  for (int i = 0; i < nh; i++) status[i] = 0;
  // one lockstep group per solve precision (FP64 factors, FP32 factors, FP32
  // factors whose preconditioner reads the half-precision copy), in that order
  int rc_all = KKT_OK;
  for (int q = 0; q < 3; q++) {
    std::vector<int> idx;
    for (int i = 0; i < nh; i++)
      if (hs[i]->solve_q(true) == q) idx.push_back(i);
    if (idx.empty()) continue;
    const int K = (int)idx.size();
    std::vector<kkt_solver*> sh(K);
    std::vector<const double*> sv(K), sb(K);
    std::vector<double*> sx(K);
    std::vector<int> sst(K, 0);
    for (int k = 0; k < K; k++) {
      sh[k] = hs[idx[k]];
      sv[k] = d_vals[idx[k]];
      sb[k] = d_b[idx[k]];
      sx[k] = d_x[idx[k]];
    }
    int rc;
    if (q == 0) rc = group_bicgstab<double, double>(g, K, sh.data(), sv.data(), sb.data(), sx.data(), tolRel, maxIter,
                                                     sst.data());
    else if (q == 1) rc = group_bicgstab<float, float>(g, K, sh.data(), sv.data(), sb.data(), sx.data(), tolRel,
                                                       maxIter, sst.data());
    else rc = group_bicgstab<float, _Float16>(g, K, sh.data(), sv.data(), sb.data(), sx.data(), tolRel, maxIter,
                                              sst.data());
    if (rc) {
      (void)hipStreamSynchronize(g->st);
      for (int k = 0; k < K; k++)
        if (sst[k] == 0) {
          sst[k] = rc;
          sh[k]->fail(rc, g->err);
        }
      if (!rc_all) rc_all = rc;
    }
    for (int k = 0; k < K; k++) status[idx[k]] = sst[k];
  }
  return rc_all;
}

int kkt_group_iterative_solve(kkt_group* g, int nh, kkt_solver* const* hs, const double* const* vals,
                              const double* const* b, double* const* x, double tolRel, int maxIter, int* status) {
  if (!g) return KKT_ERR_ARGUMENT;
  g->has_err = false;
  if (nh < 1 || nh > KKT_GROUP_MAX) return g->fail(KKT_ERR_ARGUMENT, "group size must be 1..KKT_GROUP_MAX");
  if (!hs || !vals || !b || !x || !status) return g->fail(KKT_ERR_ARGUMENT, "null argument");
  // each handle's values and right-hand side go to its own device buffers (the
  // value buffer and two solve vectors), as kkt_iterative_solve stages them;
  // then the device-array group solve, and the solutions back
  std::vector<const double*> dv(nh), db(nh);
  std::vector<double*> dx(nh);
  for (int i = 0; i < nh; i++) {
    kkt_solver* h = hs[i];
    if (!h || !vals[i] || !b[i] || !x[i]) return g->fail(KKT_ERR_ARGUMENT, "null handle or array");
    if (h->state != kkt_solver::FACTORED)
      return g->fail(KKT_ERR_STATE, "iterativeSolve called before factor (handle " + std::to_string(i) + ")");
    for (int j = 0; j < i; j++)
      if (hs[j] == h) return g->fail(KKT_ERR_ARGUMENT, "a handle appears twice in the group");
    dv[i] = h->d_vals;
    db[i] = h->vec(10);
    dx[i] = h->vec(11);
    if (hipMemcpyAsync(h->d_vals, vals[i], (size_t)h->nnz * sizeof(double), hipMemcpyHostToDevice, h->st) !=
            hipSuccess ||
        hipMemcpyAsync(h->vec(10), b[i], (size_t)h->n * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
      return g->fail(KKT_ERR_COPY, "hipMemcpyAsync (vals, b) failed");
  }
  int rc = kkt_group_iterative_solve_device(g, nh, hs, dv.data(), db.data(), dx.data(), tolRel, maxIter, status);
  for (int i = 0; i < nh; i++) {
    if (status[i] <= 0) continue;
    kkt_solver* h = hs[i];
    if (hipMemcpyAsync(x[i], dx[i], (size_t)h->n * sizeof(double), hipMemcpyDeviceToHost, g->st) != hipSuccess) {
      status[i] = h->fail(KKT_ERR_COPY, "hipMemcpyAsync (x) failed");
      if (!rc) rc = g->fail(KKT_ERR_COPY, "hipMemcpyAsync (x) failed");
    }
  }
  if (hipStreamSynchronize(g->st) != hipSuccess && !rc) rc = g->fail(KKT_ERR_DEVICE, "group stream failed");
  return rc;
}

int kkt_iterative_solve(kkt_solver* h, const double* vals, const double* b, double* x, double tolRel,
                        int maxIter) {
  CHECK_H(h);
  if (!vals || !b || !x) return h->fail(KKT_ERR_ARGUMENT, "null argument");
  if (h->state != kkt_solver::FACTORED) return h->fail(KKT_ERR_STATE, "iterativeSolve called before factor");
  if (hipMemcpyAsync(h->d_vals, vals, (size_t)h->nnz * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (vals) failed");
  double* db = h->vec(10);
  double* dx = h->vec(11);
  if (hipMemcpyAsync(db, b, (size_t)h->n * sizeof(double), hipMemcpyHostToDevice, h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (b) failed");
  int it = kkt_iterative_solve_device(h, h->d_vals, db, dx, tolRel, maxIter);
  if (it <= 0) return it;
  if (hipMemcpyAsync(x, dx, (size_t)h->n * sizeof(double), hipMemcpyDeviceToHost, h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return h->fail(KKT_ERR_COPY, "hipMemcpyAsync (x) failed");
  return it;
}

int kkt_release_factor(kkt_solver* h) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  if (h->dev_ok && h->st && hipStreamSynchronize(h->st) != hipSuccess)
    return h->fail(KKT_ERR_DEVICE, "handle stream failed");
  // fronts, packed solve arrays, workspaces and solve vectors (and the graphs
  // that name them); the analysis, its plan and the pattern stay resident
  h->release_numeric();
  h->fallback_active = false;
  h->state = kkt_solver::ANALYZED;
  h->ok();
  return KKT_OK;
}

int kkt_set_ir_steps(kkt_solver* h, int nsteps) {
  CHECK_H(h);
  if (nsteps < 0) return h->fail(KKT_ERR_ARGUMENT, "negative refinement steps");
  h->ir_steps = nsteps;
  h->ok();
  return KKT_OK;
}

const char* kkt_last_error(kkt_solver* h) {
  if (!h || !h->has_err) return nullptr;
  return h->err.c_str();
}

const char* kkt_version(void) { return "kkt_mi355x 0.1.0 (gfx950)"; }

void kkt_set_timing_enabled(int on) { g_timing = on != 0; }

int kkt_set_precision(kkt_solver* h, int prec) {
  CHECK_H(h);
  if (prec != KKT_PREC_FP64 && prec != KKT_PREC_FP32_REFINE)
    return h->fail(KKT_ERR_ARGUMENT, "unknown precision");
  h->prec = prec;
  h->ok();
  return KKT_OK;
}

int kkt_set_refine_tol(kkt_solver* h, double tol, int max_steps) {
  CHECK_H(h);
  if (!(tol > 0) || max_steps < 1) return h->fail(KKT_ERR_ARGUMENT, "bad refinement settings");
  h->refine_tol = tol;
  h->refine_max = max_steps;
  h->ok();
  return KKT_OK;
}

int kkt_set_graphs(kkt_solver* h, int on) {
  CHECK_H(h);
  h->graphs_on = on != 0;
  if (!h->graphs_on && h->dev_ok) h->drop_graphs();
  h->ok();
  return KKT_OK;
}

int kkt_set_iterative_half(kkt_solver* h, int on) {
  CHECK_H(h);
  const bool v = on != 0;
  if (v != h->half_on) {
    h->half_on = v;
    h->have16 = false;  // until the next factor
    if (h->dev_ok) h->drop_graphs();
  }
  h->ok();
  return KKT_OK;
}

int kkt_set_ordering(kkt_solver* h, int ordering) {
  CHECK_H(h);
  if (ordering != 0 && ordering != 1) return h->fail(KKT_ERR_ARGUMENT, "unknown ordering");
  h->ordering = ordering;
  h->ok();
  return KKT_OK;
}

int kkt_set_max_supernode(kkt_solver* h, int cols) {
  CHECK_H(h);
  if (cols < 64 || cols > KMAX) return h->fail(KKT_ERR_ARGUMENT, "supernode width outside 64..2048");
  h->max_super = cols;
  h->ok();
  return KKT_OK;
}

static void fill_info(const Symbolic& S, kkt_info* o) {
  o->n = S.n;
  o->nsuper = S.nsuper;
  o->nlevels = S.nlevels;
  o->ndelayed = S.ndelayed;
  o->nnzL = S.nnzL;
  o->nnzL_exact = S.nnzL_exact;
  o->factor_flops = S.flops;
}

int kkt_get_info(kkt_solver* h, kkt_info* out) {
  CHECK_H(h);
  if (!out) return KKT_ERR_ARGUMENT;
  memset(out, 0, sizeof(*out));
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  fill_info(h->S, out);
  out->nnz = h->nnz;
  out->front_bytes = h->front_elems * (long long)(h->fprec == KKT_PREC_FP64 ? 8 : 4);
  out->npos = h->ctr.npos;
  out->nneg = h->ctr.nneg;
  out->nperturbed = h->ctr.npert;
  out->n2x2 = h->ctr.n2x2;
  out->pivoted_blocks = h->ctr.npivoted;
  out->precision = h->fprec;
  out->refine_steps = h->last_refine_steps;
  out->fp64_fallbacks = h->fp64_fallbacks;
  out->last_rel_residual = h->last_relres;
  out->device_bytes = (long long)h->held_bytes();
  out->factored_bytes = (long long)(h->plan_bytes() + h->pattern_bytes() + h->numeric_bytes(h->prec));
  out->fp64_fallback_bytes =
      (h->prec == KKT_PREC_FP32_REFINE && !h->d_F64)
          ? (long long)(((size_t)h->front_elems + 2 * (size_t)h->n + 4 + (size_t)std::max<long long>(h->w_elems, 1)) * 8)
          : 0;
  out->precond_half = h->use16() ? 1 : 0;
  return KKT_OK;
}

int kkt_get_symbolic(kkt_solver* h, int* perm, int* sup_ptr, int* sparent, long long* rows_ptr, int* rows) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  const Symbolic& S = h->S;
  if (perm) memcpy(perm, S.perm.data(), sizeof(int) * S.n);
  if (sup_ptr) memcpy(sup_ptr, S.sup_ptr.data(), sizeof(int) * (S.nsuper + 1));
  if (sparent) memcpy(sparent, S.sparent.data(), sizeof(int) * S.nsuper);
  if (rows_ptr)
    for (int s = 0; s <= S.nsuper; s++) rows_ptr[s] = S.rows_ptr[s];
  if (rows && !S.rows.empty()) memcpy(rows, S.rows.data(), sizeof(int) * S.rows.size());
  return KKT_OK;
}

int kkt_verify_device_structure(kkt_solver* h) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  (void)hipStreamSynchronize(h->st);
  int bad = 0;
  bad |= cmp_dev(h->d_perm, h->S.perm);
  bad |= cmp_dev(h->ds.sn_first, h->h_first);
  bad |= cmp_dev(h->ds.sn_k, h->h_k);
  bad |= cmp_dev(h->ds.sn_m, h->h_m);
  bad |= cmp_dev(h->ds.sn_foff, h->h_foff);
  bad |= cmp_dev(h->ds.sn_roff, h->h_roff);
  bad |= cmp_dev(h->ds.ch_ptr, h->h_chp);
  bad |= cmp_dev(h->ds.rows, h->S.rows);
  bad |= cmp_dev(h->ds.rel, h->h_rel);
  return bad ? 1 : 0;
}

void kkt_timing_reset(kkt_solver* h, int enable) {
  if (!h) return;
  h->tm_on = enable != 0;
  h->tm_mask = (enable < 0) ? 0xFF : (enable & 0xFF);
  for (int c = 0; c < TC_N; c++) { h->tm_ms[c] = 0; h->tm_cnt[c] = 0; }
  h->tm_upd_flops = 0;
  h->ev_used = 0;
  h->ev_class.clear();
}

int kkt_autotune_solve(kkt_solver* h, int reps) {
  CHECK_H(h);
  int rc = h->autotune_solve(reps);
  if (!rc) h->ok();
  return rc;
}

int kkt_get_solve_shapes(kkt_solver* h, int* fwd_elems, int* bwd_elems) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  for (size_t L = 0; L < h->fr_epl.size(); L++) {
    if (fwd_elems) fwd_elems[L] = h->fr_epl[L];
    if (bwd_elems) bwd_elems[L] = h->bc_epl[L];
  }
  h->ok();
  return KKT_OK;
}

int kkt_get_forward_forms(kkt_solver* h, int* launches) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  for (size_t L = 0; L < h->fr_fuse.size(); L++)
    if (launches) launches[L] = h->fr_fuse[L] ? 1 : 2;
  h->ok();
  return KKT_OK;
}

int kkt_set_forward_forms(kkt_solver* h, const int* launches) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  if (!launches) return h->fail(KKT_ERR_ARGUMENT, "null forms");
  const int nl = (int)h->fr_fuse.size();
  for (int L = 0; L < nl; L++)
    if (launches[L] != 1 && launches[L] != 2) return h->fail(KKT_ERR_ARGUMENT, "forward form must be 1 or 2");
  for (int L = 0; L < nl; L++) h->fr_fuse[L] = launches[L] == 1 ? 1 : 0;
  int rc;
  if ((rc = h->build_solve_items())) return rc;
  h->drop_graphs();
  h->ok();
  return KKT_OK;
}

int kkt_time_kernels(kkt_solver* h, int what, int reps, double* ms_per_launch, int* launches_per_rep) {
  CHECK_H(h);
  if (!ms_per_launch || !launches_per_rep) return h->fail(KKT_ERR_ARGUMENT, "null output");
  int rc = h->time_kernels(what, reps, ms_per_launch, launches_per_rep);
  if (!rc) h->ok();
  return rc;
}

int kkt_get_forward_row_groups(kkt_solver* h, int* groups) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  if (!groups) return h->fail(KKT_ERR_ARGUMENT, "null groups");
  for (size_t L = 0; L < h->fr_rg.size(); L++) groups[L] = h->fr_rg[L];
  h->ok();
  return KKT_OK;
}

int kkt_set_forward_row_groups(kkt_solver* h, const int* groups) {
  CHECK_H(h);
  if (h->state < kkt_solver::ANALYZED) return h->fail(KKT_ERR_STATE, "no analysis");
  if (!groups) return h->fail(KKT_ERR_ARGUMENT, "null groups");
  const int nl = (int)h->fr_rg.size();
  for (int L = 0; L < nl; L++)
    if (groups[L] != 1 && groups[L] != 2 && groups[L] != 4 && groups[L] != 8)
      return h->fail(KKT_ERR_ARGUMENT, "row groups must be 1, 2, 4 or 8");
  for (int L = 0; L < nl; L++) h->fr_rg[L] = groups[L];
  int rc;
  if ((rc = h->build_solve_items())) return rc;
  h->drop_graphs();
  h->ok();
  return KKT_OK;
}

int kkt_timing_get(kkt_solver* h, double* ms, long long* launches, double* update_flops) {
  CHECK_H(h);
  for (int c = 0; c < TC_N; c++) {
    if (ms) ms[c] = h->tm_ms[c];
    if (launches) launches[c] = h->tm_cnt[c];
  }
  if (update_flops) *update_flops = h->tm_upd_flops;
  return KKT_OK;
}

// ---- host-only symbolic analysis ----
struct kkt_symbolic {
  Symbolic S;
};

kkt_symbolic* kkt_symbolic_create(int n, int nnz, const int* rowOffs, const int* colIdxs,
                                  const unsigned char* delay, int ordering) {
  std::string e;
  if (validate_pattern(n, nnz, rowOffs, colIdxs, e) != 0) return nullptr;
  kkt_symbolic* s = new kkt_symbolic();
  SymOptions opt;
  opt.ordering = ordering;
  if (symbolic_analyze(n, rowOffs, colIdxs, delay, opt, s->S, e) != 0) {
    delete s;
    return nullptr;
  }
  return s;
}

int kkt_symbolic_info(kkt_symbolic* s, kkt_info* out) {
  if (!s || !out) return KKT_ERR_ARGUMENT;
  memset(out, 0, sizeof(*out));
  fill_info(s->S, out);
  return KKT_OK;
}

int kkt_symbolic_get(kkt_symbolic* s, int* perm, int* parent, int* colcount, int* sup_ptr, int* sparent,
                     long long* rows_ptr, int* rows, int* level) {
  if (!s) return KKT_ERR_ARGUMENT;
  const Symbolic& S = s->S;
  if (perm) memcpy(perm, S.perm.data(), sizeof(int) * S.n);
  if (parent) memcpy(parent, S.parent.data(), sizeof(int) * S.n);
  if (colcount) memcpy(colcount, S.colcount.data(), sizeof(int) * S.n);
  if (sup_ptr) memcpy(sup_ptr, S.sup_ptr.data(), sizeof(int) * (S.nsuper + 1));
  if (sparent) memcpy(sparent, S.sparent.data(), sizeof(int) * S.nsuper);
  if (rows_ptr)
    for (int q = 0; q <= S.nsuper; q++) rows_ptr[q] = S.rows_ptr[q];
  if (rows && !S.rows.empty()) memcpy(rows, S.rows.data(), sizeof(int) * S.rows.size());
  if (level) memcpy(level, S.level.data(), sizeof(int) * S.nsuper);
  return KKT_OK;
}

void kkt_symbolic_destroy(kkt_symbolic* s) { delete s; }

// ---- batched systems ----
#define CHECK_B(b) \
  if (!(b)) return KKT_ERR_ARGUMENT
kkt_batch* kkt_batch_create(void) {
  try {
    return new kkt_batch();
  } catch (...) {
    return nullptr;
  }
}

void kkt_batch_destroy(kkt_batch* b) { delete b; }

int kkt_batch_set_systems(kkt_batch* b, int nsys, const int* sysN, const int* rowOffs, const int* colIdxs,
                          int mtype) {
  CHECK_B(b);
  int rc = b->set_systems(nsys, sysN, rowOffs, colIdxs, mtype);
  if (!rc) b->has_err = false;
  return rc;
}

int kkt_batch_analyze(kkt_batch* b, const double* vals) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "analyze called before setSystems");
  b->free_dev();
  return b->from_h(vals ? kkt_analyze_values(b->h, vals) : kkt_analyze(b->h));
}

int kkt_batch_factor(kkt_batch* b, const double* vals) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "factor called before setSystems");
  return b->from_h(kkt_factor(b->h, vals));
}

int kkt_batch_factor_device(kkt_batch* b, const double* d_vals) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "factor called before setSystems");
  return b->from_h(kkt_factor_device(b->h, d_vals));
}

int kkt_batch_solve_device(kkt_batch* b, const double* d_B, double* d_X) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "solve called before setSystems");
  if (!d_B || !d_X || d_B == d_X) return b->fail(KKT_ERR_ARGUMENT, "null or aliased vectors");
  return b->solve_device(d_B, d_X);
}

int kkt_batch_solve(kkt_batch* b, const double* B, double* X) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "solve called before setSystems");
  if (!B || !X) return b->fail(KKT_ERR_ARGUMENT, "null vector");
  kkt_solver* h = b->h;
  if (h->state != kkt_solver::FACTORED) return b->fail(KKT_ERR_STATE, "solve called before factor");
  double* db = h->vec(10);
  double* dx = h->vec(11);
  const size_t bytes = (size_t)h->n * sizeof(double);
  if (hipMemcpyAsync(db, B, bytes, hipMemcpyHostToDevice, h->st) != hipSuccess)
    return b->fail(KKT_ERR_COPY, "hipMemcpyAsync (B) failed");
  int rc = b->solve_device(db, dx);
  if (rc) return rc;
  if (hipMemcpyAsync(X, dx, bytes, hipMemcpyDeviceToHost, h->st) != hipSuccess ||
      hipStreamSynchronize(h->st) != hipSuccess)
    return b->fail(KKT_ERR_COPY, "hipMemcpyAsync (X) failed");
  return KKT_OK;
}

int kkt_batch_set_precision(kkt_batch* b, int prec) {
  CHECK_B(b);
  if (!b->h && !(b->h = kkt_create())) return b->fail(KKT_ERR_INIT, "handle creation failed");
  return b->from_h(kkt_set_precision(b->h, prec));
}

int kkt_batch_set_refine_tol(kkt_batch* b, double tol, int max_steps) {
  CHECK_B(b);
  if (!b->h && !(b->h = kkt_create())) return b->fail(KKT_ERR_INIT, "handle creation failed");
  return b->from_h(kkt_set_refine_tol(b->h, tol, max_steps));
}

int kkt_batch_set_ir_steps(kkt_batch* b, int nsteps) {
  CHECK_B(b);
  if (!b->h && !(b->h = kkt_create())) return b->fail(KKT_ERR_INIT, "handle creation failed");
  return b->from_h(kkt_set_ir_steps(b->h, nsteps));
}

int kkt_batch_autotune_solve(kkt_batch* b, int reps) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "no systems");
  return b->from_h(kkt_autotune_solve(b->h, reps));
}

int kkt_batch_get_info(kkt_batch* b, kkt_info* out) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "no systems");
  return b->from_h(kkt_get_info(b->h, out));
}

int kkt_batch_system_info(kkt_batch* b, int sys, kkt_batch_sysinfo* out) {
  CHECK_B(b);
  if (!b->h) return b->fail(KKT_ERR_STATE, "no systems");
  return b->system_info(sys, out);
}

const char* kkt_batch_last_error(kkt_batch* b) {
  if (!b || !b->has_err) return nullptr;
  return b->err.c_str();
}

}  // extern "C"
