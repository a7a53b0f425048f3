// Device numeric factorization: the symbolic plan is compiled once into a
// program of batched kernel launches (per elimination-tree level), replayed
// for every numeric factorization of the same pattern.
//
// Per level (all supernodes of equal height together):
//   1. assemble fronts: zero, scatter A, extend-add children updates
//   2. uncompressed supernodes (batched over the level): blocked Cholesky of
//      the (ncols+R) x ncols panel, then update -= L21 L21^T
//   3. compressed separators: in-place HODLR factor of the diagonal in the
//      reference's traversal order (FactorManager.cpp:8907-8992), explicit
//      blocks factored densely then compressed (DiagonalBlock.cpp:393-453,
//      1384-1615), recursive off-diagonal blocks compressed from
//      S L_left^{-T} (FactorManager.cpp:8998-9036); then L21 = F21 L11^{-T},
//      the randomized range finder with power iteration on L21
//      (FactorManager.cpp:8193-8297, 8672-8779), V = L21 U
//      (FactorManager.cpp:5691-5751), and update -= V (U^T U) V^T
//      (Supernode.cpp:1266-1429).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <functional>
#include <stdexcept>
#include <thread>

#include "factor.h"
#include "solve.h"

namespace rsc {

namespace {

enum OpKind {
  OP_ASM = 0,
  OP_GEMM_NN,
  OP_GEMM_TN,
  OP_GEMM_NT,
  OP_GEMM_TT,
  OP_POTRF,
  OP_TRSM,
  OP_COPY,
  OP_SHIFT,
  OP_EYE,
  OP_CHECK,
  OP_CONV
};

struct Step {
  int op;
  int count;
  int64_t first;
  int f64 = 0;   // 1: run the op in FP64 on FP64 buffers (QR of an FP32 factor)
  int bf16 = 0;  // 1: GEMM products in BF16 with FP32 accumulation (sketches)
};

// A step recorded while the compressed supernodes of one level are captured
// (one list per supernode); lists are then merged into shared launches.
struct Pending {
  int kind;  // OpKind (the GEMM variants carry their transposes)
  bool lower = false;
  int f64 = 0, bf16 = 0;
  std::string ctx;
  std::vector<dev::GemmJob> g;
  std::vector<dev::PotrfJob> p;
  std::vector<dev::TrsmJob> t;
  std::vector<dev::CopyJob> c;
  std::vector<dev::EyeJob> e;
  std::vector<dev::ConvJob> v;
  std::vector<dev::ShiftJob> s;
};

struct CholMat {
  char* A;
  int n, m, ld, node;
  int tolerant;
};
struct TrsmMat {
  char* X;
  int m, ldx;
  const char* L;
  int n, ldl;
};
struct SketchJob {
  char* X;
  int rows, cols, ldx;
  const char* G;  // r x rows column-major (ld r)
  int r;
  char* U;  // cols x r, ld cols (also holds Y)
  char* V;  // rows x r, ld rows
  char* Z;  // rows x r
  char* W;  // r x r
  int node;
};

// Tree node q lies in the subtree rooted at sub (q == sub included).
inline bool inside_subtree(const Supernode& s, int q, int sub) {
  for (int a = q; a >= 0; a = s.hnodes[a].parent)
    if (a == sub) return true;
  return false;
}

// One dataflow task list on the device.
struct DfList {
  DBuf tasks, gv, gw, cnt;
  // ready-queue dispatch: queue slots, pop/push counters, pending waits per
  // group and their initial values, tasks per group, successor groups
  DBuf q_queue, q_ctr, q_pend, q_pend0, q_need, q_gptr, q_gtasks, q_ginfo, q_sptr, q_succ, q_init;
  int q_n0 = 0;
  int ntasks = 0, ngroups = 0, grid = 1;
  // kept for the optional timing trace
  std::vector<dev::DfTask> host;
  std::vector<dev::DfWait> host_wait;
  std::vector<dev::GvTask> host_gv;
  std::vector<float> host_feat;
};

template <typename T>
class Engine final : public FactorEngine {
 public:
  Engine(const Plan& plan, const Csc& A, hipStream_t st)
      : plan_(plan), A_(A), st_(st) {
    try {
      start_rng();
      setup_static();
    } catch (...) {
      // the helper threads already started are joined before the members
      // go (a joinable std::thread would end the process)
      shutdown();
      throw;
    }
  }
  ~Engine() override { shutdown(); }
  void shutdown() {
    if (prog_thread_.joinable()) prog_thread_.join();
    (void)hipStreamSynchronize(st_);
    if (rng_thread_.joinable()) rng_thread_.join();
    if (df_thread_.joinable()) df_thread_.join();
    if (rng_host_) (void)hipHostFree(rng_host_);
    rng_host_ = nullptr;
    drop_graph();
  }

  void numeric(const double* Ax) override;
  void apply(int k, const double* B, int ldb, double* X, int ldx, hipStream_t s,
             const int* live, int multi_min_k) override;
  void check_apply(hipStream_t s) override;
  int64_t apply_scratch_doubles(int k) const override;
  void apply_into(int k, const double* B, int ldb, double* X, int ldx, hipStream_t s,
                  double* scratch) override;
  const int* apply_status_device() const override { return dapst_.as<int>(); }
  void raise_apply_status(const int* words) override;
  int64_t factor_bytes() const override { return (int64_t)factor_.bytes; }
  int64_t workspace_bytes() const override { return (int64_t)work_.bytes; }
  int pivot_replacements() const override { return pivot_replacements_; }
  void release_workspace() override {
    if (prog_thread_.joinable()) prog_thread_.join();
    prog_err_ = nullptr;
    (void)hipStreamSynchronize(st_);
    drop_graph();
    work_.reset();
    program_ready_ = false;
  }

 private:
  const Plan& plan_;
  const Csc& A_;
  hipStream_t st_;
  // Gaussian sketch entries: drawn on host threads into pinned memory while
  // the host plans and the device runs the levels below the first compressed
  // supernode, then copied once before the first step that reads them.
  std::thread rng_thread_;
  T* rng_host_ = nullptr;
  std::exception_ptr rng_err_;
  int rng_step_ = -1;  // first program step reading the sketch entries
  bool rng_on_device_ = false;
  void start_rng() {
    if (plan_.rng_count <= 0 || rng_host_ || rng_thread_.joinable()) return;
    rng_thread_ = std::thread([this]() {
      try {
        void* p = nullptr;
        RSC_HIP(hipHostMalloc(&p, (size_t)plan_.rng_count * sizeof(T)));
        gaussian_stream(plan_.opt.seed, plan_.rng_count, (T*)p);
        rng_host_ = (T*)p;
      } catch (...) {
        rng_err_ = std::current_exception();
      }
    });
  }
  void upload_rng() {
    // a compressed supernode may draw nothing (a single diagonal leaf and no
    // off-diagonal rows): then there is no stream to copy
    if (plan_.rng_count <= 0) {
      rng_on_device_ = true;
      return;
    }
    start_rng();
    if (rng_thread_.joinable()) rng_thread_.join();
    if (rng_err_) {
      std::exception_ptr e = rng_err_;
      rng_err_ = nullptr;
      std::rethrow_exception(e);
    }
    if (!rng_host_) throw std::runtime_error("sketch entries were not drawn");
    RSC_HIP(hipMemcpyAsync(Wb() + rng_base(), rng_host_,
                           (size_t)plan_.rng_count * sizeof(T), hipMemcpyHostToDevice,
                           st_));
    rng_on_device_ = true;
  }
  DBuf factor_, work_, dAx_, dlowp_, da_dst_, da_src_, dfronts_, dchildren_,
      drelmap_, dtasks_, dgjobs_, dgtiles_, dpjobs_, dtjobs_, dttiles_, dcjobs_,
      dctiles_, dsjobs_, dstatus_, dejobs_, detiles_, dkjobs_, dvjobs_;
  DBuf dvnodes_, dgv_, drows_, dlevnodes_, dperm_, dpinv_, dfv_, dtv_, dw_, dz_, dranges_;
  std::vector<std::vector<SolvePhase>> fwd_, bwd_;
  std::vector<std::vector<int>> level_nodes_;  // per level, nodes outside subtrees
  int64_t tv_stride_ = 0, ntasks_ = 0, part_stride_ = 1;
  DBuf dpart_;
  std::vector<int64_t> fv_off_;
  // dataflow apply (one persistent launch per application)
  DBuf dapst_, dtrace_;
  DBuf dstptr_, dstnodes_, dstcols_, dpanel_;
  DBuf dmsptr_, dmsnodes_, dmscols_;  // subtrees of the multi-right-hand-side list
  int nsub_ = 0;
  // multi: the subtrees of the multi-right-hand-side list
  dev::SubtreeArgs subtree_args(const double* B, int ldb, bool multi = false) const {
    dev::SubtreeArgs a{};
    a.ptr = (multi ? dmsptr_ : dstptr_).as<int>();
    a.nodes = (multi ? dmsnodes_ : dstnodes_).as<int>();
    a.cols = (multi ? dmscols_ : dstcols_).as<int>();
    a.panel = dpanel_.as<int64_t>();
    a.vn = dvnodes_.as<dev::VecNode>();
    a.children = dchildren_.as<int>();
    a.relmap = drelmap_.as<int>();
    a.rows = drows_.as<int>();
    a.perm = dperm_.as<int>();
    a.B = B;
    a.ldb = ldb;
    return a;
  }
  SolveLayout sl_;
  // dataflow task lists: [0] with subtree tasks (single right-hand side),
  // [1] without (multi-right-hand-side apply)
  DfList df_[2];
  int df_last_ = 0;  // list of the last apply (for the trace)
  std::thread df_thread_;
  std::unique_ptr<DataflowTasks> df_pending_, df_pending2_;
  std::exception_ptr df_err_;
  bool solve_ready_ = false;
  // longest pause between polls of a dataflow wait or an empty queue slot,
  // in s_sleep(4) units: short pauses see a pushed task sooner (measured
  // best at 1-4 for 64^3 and 2-4 for 128^3; 16 was 2-3% slower)
  int df_backoff_ = 2;
  int df_resident_[2] = {1, 1};  // resident workgroups: single- / multi-RHS kernel
  void finish_solve_setup();
  std::string df_trace_path_;

  unsigned long long df_timeout_ = 0;
  bool use_dataflow_ = false;
  bool df_queue_ = true;  // ready-queue dispatch (RSC_DF_QUEUE=0: tickets)
  int64_t fv_stride_ = 0;
  int fv_k_ = 0;
  std::vector<std::vector<int>> levnode_off_;  // per level: offset in dlevnodes
  std::vector<int> lev_off_;
  std::vector<int> lev_front_;  // largest front (ncols + R) per level
  bool program_ready_ = false;
  // The program captured as a graph (kernel nodes only), replayed by later
  // numeric factorizations of the same pattern; rebuilt with the program.
  hipGraph_t fgraph_ = nullptr;
  hipGraphExec_t fgexec_ = nullptr;
  bool graph_failed_ = false;
  void drop_graph() {
    if (fgexec_) (void)hipGraphExecDestroy(fgexec_);
    if (fgraph_) (void)hipGraphDestroy(fgraph_);
    fgexec_ = nullptr;
    fgraph_ = nullptr;
  }
  void run_program(std::vector<hipEvent_t>* ev);
  // the first program build runs on this thread during the engine setup
  std::thread prog_thread_;
  std::exception_ptr prog_err_;
  void join_program_build() {
    if (prog_thread_.joinable()) prog_thread_.join();
    if (prog_err_) {
      std::exception_ptr e = prog_err_;
      prog_err_ = nullptr;
      program_ready_ = false;
      std::rethrow_exception(e);
    }
  }
  int pivot_replacements_ = 0;
  // program
  std::vector<Step> prog_;
  std::vector<dev::GemmJob> gjobs_;
  std::vector<dev::GemmTile> gtiles_;
  int ngcnt_ = 0;            // split-k tile counters of the program
  int64_t gpart_bytes_ = 0;  // largest split-k partial set of one launch
  DBuf dgpart_, dgcnt_;
  std::vector<dev::PotrfJob> pjobs_;
  std::vector<dev::TrsmJob> tjobs_;
  std::vector<dev::RowTile> ttiles_;
  std::vector<dev::CopyJob> cjobs_;
  std::vector<dev::RowTile> ctiles_;
  std::vector<dev::ShiftJob> sjobs_;
  std::vector<dev::EyeJob> ejobs_;
  std::vector<dev::CheckJob> kjobs_;
  bool debug_check_ = false;
  std::vector<std::string> step_desc_;
  std::vector<std::string> step_ctx_;
  std::string ctx_;
  int cur_f64_ = 0;
  int cur_bf16_ = 0;  // set around the range-finder products of BF16 sketches
  bool bf16_sketch() const {
    return sizeof(T) == 4 && plan_.opt.sketch_precision == RSC_BF16;
  }
  int64_t esize() const { return cur_f64_ ? 8 : (int64_t)sizeof(T); }
  std::vector<dev::ConvJob> vjobs_;
  int64_t qr64_elems_ = 0;  // FP64 QR scratch (in T elements)
  int64_t zp_elems_ = 0;    // Zp = L21 Lc^{-1} staging of one level (T elements)
  // Solve-helper targets of a compressed diagonal: matrices S (rows of
  // subtree `sub`, r columns) that receive D = L_sub^{-1} S (W, from V) or
  // D = L_sub^{-T} S (X, from U).  Internal tree nodes give one W and one X
  // target; the off-diagonal pair gives X_off = L_D^{-T} U (subtree = root).
  struct HTarget {
    int sub;
    bool isW;
    int64_t src, dst;  // factor offsets
    int rank;
  };
  std::vector<HTarget> helper_targets(int j) const;
  void push(const Step& st0) {
    Step st = st0;
    st.f64 = cur_f64_;
    st.bf16 = cur_bf16_;
    prog_.push_back(st);
    step_ctx_.push_back(ctx_);
  }
  void add_checks();
  void write_factor_trace(const char* path, std::vector<hipEvent_t>& ev);
  std::vector<dev::RowTile> etiles_;
  std::vector<dev::AsmTask> tasks_;
  int64_t inv_elems_ = 0;
  int64_t scratch_elems_ = 0;
  std::vector<int64_t> node_scratch_;   // per supernode
  std::vector<std::vector<int64_t>> hnode_scratch_;

  // FP32 factors replace failed pivots instead of reporting them
  static int factor_tol() { return sizeof(T) == 4 ? 2 : 0; }
  T* F() const { return factor_.as<T>(); }
  T* Wb() const { return work_.as<T>(); }
  char* fp(int64_t off) const { return (char*)(F() + off); }
  char* wp(int64_t off) const { return (char*)(Wb() + off); }
  int64_t upd_base() const { return plan_.work_elems; }
  int64_t scratch_base() const { return plan_.work_elems + plan_.update_elems; }
  int64_t rng_base() const { return scratch_base() + scratch_elems_; }
  int64_t inv_base() const { return rng_base() + plan_.rng_count; }
  int64_t qr64_base() const { return inv_base() + inv_elems_; }
  int64_t zp_base() const { return qr64_base() + qr64_elems_; }

  void setup_static();
  void build_program();
  void layout_scratch();

  // emission helpers
  void emit_gemm(int op, const std::vector<dev::GemmJob>& jobs, bool lower);
  void emit_potrf(const std::vector<dev::PotrfJob>& jobs);
  void emit_trsm(const std::vector<dev::TrsmJob>& jobs);
  void emit_copy(const std::vector<dev::CopyJob>& jobs);
  void emit_eye(const std::vector<dev::EyeJob>& jobs);
  void emit_conv(const std::vector<dev::ConvJob>& jobs);
  void emit_shift(const std::vector<dev::ShiftJob>& jobs);
  // Cross-node batching: while capture_ is set the emitters record into it
  std::vector<Pending>* capture_ = nullptr;
  Pending& record(int kind) {
    capture_->emplace_back();
    Pending& p = capture_->back();
    p.kind = kind;
    p.f64 = cur_f64_;
    p.bf16 = cur_bf16_;
    p.ctx = ctx_;
    return p;
  }
  void emit_level_compressed(const std::vector<int>& nodes);
  void run_apply(int k, const double* B, int ldb, double* X, int ldx, hipStream_t s,
                 const dev::VecBufs& vb);
  int64_t inv_cursor_ = 0, qr64_cursor_ = 0;  // scratch of the supernode being emitted
  std::vector<int64_t> inv_need_, qr64_need_;  // per supernode (T elements)
  void invert(const std::vector<CholMat>& tris);
  void panel_chol(const std::vector<CholMat>& mats);
  void trsm_right(const std::vector<TrsmMat>& mats);
  void qr(const std::vector<struct QrMat>& jobs);
  void sketch(const std::vector<SketchJob>& jobs);
  void lowrank_update(char* C, int ldc, int rows, const char* V, const char* U,
                      int cols, int r, char* K, char* Tm);
  void hsolve_right(const Supernode& s, int q, char* X, int m, int ldx, char* Tm);
  void emit_compressed(int j);
  void emit_solve_helpers(int j);
  std::vector<int64_t> wx_scratch_, wx_elems_;
};

template <typename T>
void Engine<T>::emit_gemm(int op, const std::vector<dev::GemmJob>& jobs,
                          bool lower) {
  if (capture_) {
    if (!jobs.empty()) {
      Pending& p = record(op);
      p.lower = lower;
      p.g = jobs;
    }
    return;
  }
  Step st{op, 0, (int64_t)gtiles_.size()};
  const int base = (int)gjobs_.size();
  gjobs_.insert(gjobs_.end(), jobs.begin(), jobs.end());
  const size_t t0 = gtiles_.size();
  const int parts = dev::plan_gemm_tiles(gjobs_.data() + base, (int)jobs.size(), lower,
                                         gtiles_, &ngcnt_);
  for (size_t t = t0; t < gtiles_.size(); ++t) gtiles_[t].job += base;
  gpart_bytes_ = std::max<int64_t>(gpart_bytes_,
                                   (int64_t)parts * dev::TILE * dev::TILE * esize());
  st.count = (int)(gtiles_.size() - st.first);
  if (st.count > 0) push(st);
}

template <typename T>
void Engine<T>::emit_potrf(const std::vector<dev::PotrfJob>& jobs) {
  if (jobs.empty()) return;
  if (capture_) {
    record(OP_POTRF).p = jobs;
    return;
  }
  Step st{OP_POTRF, (int)jobs.size(), (int64_t)pjobs_.size()};
  pjobs_.insert(pjobs_.end(), jobs.begin(), jobs.end());
  push(st);
}

template <typename T>
void Engine<T>::emit_trsm(const std::vector<dev::TrsmJob>& jobs) {
  if (capture_) {
    if (!jobs.empty()) record(OP_TRSM).t = jobs;
    return;
  }
  Step st{OP_TRSM, 0, (int64_t)ttiles_.size()};
  for (const auto& J : jobs) {
    if (J.m <= 0) continue;
    int jid = (int)tjobs_.size();
    tjobs_.push_back(J);
    for (int r = 0; r < J.m; r += dev::TILE) ttiles_.push_back({jid, r});
  }
  st.count = (int)(ttiles_.size() - st.first);
  if (st.count > 0) push(st);
}

template <typename T>
void Engine<T>::emit_copy(const std::vector<dev::CopyJob>& jobs) {
  if (capture_) {
    if (!jobs.empty()) record(OP_COPY).c = jobs;
    return;
  }
  Step st{OP_COPY, 0, (int64_t)ctiles_.size()};
  for (const auto& J : jobs) {
    if (J.m <= 0 || J.n <= 0) continue;
    int jid = (int)cjobs_.size();
    cjobs_.push_back(J);
    for (int c = 0; c < J.n; c += dev::TILE) ctiles_.push_back({jid, c});
  }
  st.count = (int)(ctiles_.size() - st.first);
  if (st.count > 0) push(st);
}

// Replace each n x n lower triangle L (leading dimension ld) by L^{-T}
// (upper): X = I, X <- X L^{-T} (blocked triangular solve), copy back.  The
// solves then apply diagonal blocks as matrix-vector products.
template <typename T>
void Engine<T>::emit_eye(const std::vector<dev::EyeJob>& jobs) {
  if (jobs.empty()) return;
  if (capture_) {
    record(OP_EYE).e = jobs;
    return;
  }
  Step st{OP_EYE, 0, (int64_t)etiles_.size()};
  for (auto& J : jobs) {
    int jid = (int)ejobs_.size();
    ejobs_.push_back(J);
    for (int c = 0; c < J.n; c += dev::TILE) etiles_.push_back({jid, c});
  }
  st.count = (int)(etiles_.size() - st.first);
  if (st.count > 0) push(st);
}

template <typename T>
void Engine<T>::emit_conv(const std::vector<dev::ConvJob>& jobs) {
  if (jobs.empty()) return;
  if (capture_) {
    record(OP_CONV).v = jobs;
    return;
  }
  Step st{OP_CONV, (int)jobs.size(), (int64_t)vjobs_.size()};
  vjobs_.insert(vjobs_.end(), jobs.begin(), jobs.end());
  push(st);
}

template <typename T>
void Engine<T>::emit_shift(const std::vector<dev::ShiftJob>& jobs) {
  if (jobs.empty()) return;
  if (capture_) {
    record(OP_SHIFT).s = jobs;
    return;
  }
  Step st{OP_SHIFT, (int)jobs.size(), (int64_t)sjobs_.size()};
  sjobs_.insert(sjobs_.end(), jobs.begin(), jobs.end());
  push(st);
}

template <typename T>
void Engine<T>::invert(const std::vector<CholMat>& tris) {
  if (tris.empty()) return;
  std::vector<dev::EyeJob> ej;
  std::vector<TrsmMat> tm;
  std::vector<dev::CopyJob> cp;
  int64_t off = inv_cursor_;
  for (auto& t : tris) {
    T* X = Wb() + off;
    off += (int64_t)t.n * t.n;
    ej.push_back({X, t.n, t.n});
    tm.push_back({(char*)X, t.n, t.n, t.A, t.n, t.ld});
    cp.push_back({X, t.A, t.n, t.n, t.n, t.ld});
  }
  if (off > inv_base() + inv_elems_) throw std::runtime_error("inverse scratch too small");
  emit_eye(ej);
  trsm_right(tm);
  emit_copy(cp);
}

// Blocked right-looking Cholesky of the leading n columns of an m x n panel
// (m >= n): the trailing rows below n receive the triangular solve.
template <typename T>
void Engine<T>::panel_chol(const std::vector<CholMat>& mats) {
  const int NB = dev::TILE;
  int maxn = 0;
  for (auto& M : mats) maxn = std::max(maxn, M.n);
  for (int s0 = 0; s0 < maxn; s0 += NB) {
    std::vector<dev::PotrfJob> pj;
    std::vector<dev::TrsmJob> tj;
    std::vector<dev::GemmJob> gj;
    const int64_t es = esize();
    for (auto& M : mats) {
      if (M.n <= s0) continue;
      const int nb = std::min(NB, M.n - s0);
      char* A = M.A;
      auto at = [&](int64_t i, int64_t j) { return A + (i + j * M.ld) * es; };
      char* D = at(s0, s0);
      pj.push_back({D, nb, M.ld, M.node, M.tolerant});
      const int below = M.m - s0 - nb;
      if (below > 0) tj.push_back({D, at(s0 + nb, s0), below, nb, M.ld, M.ld});
      const int s1 = s0 + nb;
      if (M.n > s1) {
        dev::GemmJob g;
        g.A = at(s1, s0);
        g.B = at(s1, s0);
        g.C = at(s1, s1);
        g.m = M.m - s1;
        g.n = M.n - s1;
        g.k = nb;
        g.lda = g.ldb = g.ldc = M.ld;
        g.alpha = -1.0;
        g.beta = 1.0;
        gj.push_back(g);
      }
    }
    emit_potrf(pj);
    emit_trsm(tj);
    emit_gemm(OP_GEMM_NT, gj, true);
  }
}

// X(m x n) <- X L^{-T}, L n x n lower, blocked by TILE columns.
template <typename T>
void Engine<T>::trsm_right(const std::vector<TrsmMat>& mats) {
  const int NB = dev::TILE;
  int maxn = 0;
  for (auto& M : mats) maxn = std::max(maxn, M.n);
  for (int s0 = 0; s0 < maxn; s0 += NB) {
    std::vector<dev::TrsmJob> tj;
    std::vector<dev::GemmJob> gj;
    const int64_t es = esize();
    for (auto& M : mats) {
      if (M.n <= s0 || M.m <= 0) continue;
      const int nb = std::min(NB, M.n - s0);
      char* X = M.X;
      const char* L = M.L;
      tj.push_back({L + (s0 + (int64_t)s0 * M.ldl) * es, X + (int64_t)s0 * M.ldx * es, M.m,
                    nb, M.ldl, M.ldx});
      const int s1 = s0 + nb;
      if (M.n > s1) {
        dev::GemmJob g;
        g.A = X + (int64_t)s0 * M.ldx * es;
        g.B = L + (s1 + (int64_t)s0 * M.ldl) * es;
        g.C = X + (int64_t)s1 * M.ldx * es;
        g.m = M.m;
        g.n = M.n - s1;
        g.k = nb;
        g.lda = M.ldx;
        g.ldb = M.ldl;
        g.ldc = M.ldx;
        g.alpha = -1.0;
        g.beta = 1.0;
        gj.push_back(g);
      }
    }
    emit_trsm(tj);
    emit_gemm(OP_GEMM_NT, gj, false);
  }
}

// Thin QR by shifted CholeskyQR3 (Fukaya et al. 2020: one shifted pass, then
// CholeskyQR2): each Y (m x r) becomes an orthonormal basis of its range.
// A column that is numerically dependent on the previous ones (a rank-
// deficient sketch: the block has fewer directions than the rank asks for)
// comes out zero instead of as amplified rounding noise, so U U^T is always
// an orthogonal projector (of lower rank there), which keeps every
// compressed update within the exact one and the factor SPD.
struct QrMat {
  char* Y;
  int m, r;
  char* W;  // r x r scratch
  int node;
};

template <typename T>
void Engine<T>::qr(const std::vector<QrMat>& jobs0) {
  std::vector<QrMat> jobs = jobs0;
  const bool to64 = sizeof(T) == 4;
  if (to64) {
    // FP32 factor: orthonormalize in FP64 (CholeskyQR's guarantees need
    // 11 (m r + r (r+1)) u < 1, which FP32 does not give at these sizes)
    double* base = (double*)(Wb() + qr64_cursor_);
    int64_t off = 0;
    std::vector<dev::ConvJob> cv;
    for (auto& J : jobs) {
      double* Y64 = base + off;
      off += (int64_t)J.m * J.r;
      double* W64 = base + off;
      off += (int64_t)J.r * J.r;
      cv.push_back({J.Y, Y64, (int64_t)J.m * J.r, 1, 0});
      J.Y = (char*)Y64;
      J.W = (char*)W64;
    }
    if (qr64_cursor_ - qr64_base() + 2 * off > qr64_elems_)
      throw std::runtime_error("qr64 scratch too small");
    emit_conv(cv);
    cur_f64_ = 1;
  }
  for (int pass = 0; pass < 3; ++pass) {
    std::vector<dev::GemmJob> gram;
    std::vector<dev::ShiftJob> sh;
    std::vector<CholMat> ch;
    std::vector<TrsmMat> tr;
    for (auto& J : jobs) {
      dev::GemmJob g;
      g.A = J.Y;
      g.B = J.Y;
      g.C = J.W;
      g.m = J.r;
      g.n = J.r;
      g.k = J.m;
      g.lda = g.ldb = J.m;
      g.ldc = J.r;
      g.alpha = 1.0;
      g.beta = 0.0;
      gram.push_back(g);
      sh.push_back({J.W, J.r, J.r, J.m, 0});
      // pass 0 is shifted (pivots stay positive); passes 1 and 2 see the
      // Gram of columns of norm <= 1 and drop dependent ones
      ch.push_back({J.W, J.r, J.r, J.r, J.node, pass == 0 ? 1 : 3});
      tr.push_back({J.Y, J.m, J.m, J.W, J.r, J.r});
    }
    emit_gemm(OP_GEMM_TN, gram, false);
    if (pass == 0) emit_shift(sh);
    panel_chol(ch);
    trsm_right(tr);
  }
  if (to64) {
    cur_f64_ = 0;
    std::vector<dev::ConvJob> cv;
    for (size_t q = 0; q < jobs.size(); ++q)
      cv.push_back({jobs[q].Y, jobs0[q].Y, (int64_t)jobs[q].m * jobs[q].r, 0, 0});
    emit_conv(cv);
  }
}

// Randomized range finder on X (rows x cols): Y = X^T G, power iterations
// Y = X^T (X Y), orthonormal basis U of range(Y), V = X U.  Every product is
// orthonormalized before the next multiply (subspace iteration): the range
// is that of the un-normalized iteration, but each QR sees cond(X) instead
// of cond(X)^3.
template <typename T>
void Engine<T>::sketch(const std::vector<SketchJob>& jobs) {
  std::vector<dev::GemmJob> g0, g1, g2, g3;
  std::vector<QrMat> qu, qz;
  for (auto& J : jobs) {
    dev::GemmJob g;
    g.A = J.X;
    g.B = J.G;
    g.C = J.U;
    g.m = J.cols;
    g.n = J.r;
    g.k = J.rows;
    g.lda = J.ldx;
    g.ldb = J.r;
    g.ldc = J.cols;
    g.alpha = 1.0;
    g.beta = 0.0;
    g0.push_back(g);
    qu.push_back({J.U, J.cols, J.r, J.W, J.node});
    qz.push_back({J.Z, J.rows, J.r, J.W, J.node});
  }
  // range-finder products (BF16 with FP32 accumulation for BF16 sketches);
  // V = X U at the end stays in the factor precision
  cur_bf16_ = bf16_sketch() ? 1 : 0;
  emit_gemm(OP_GEMM_TT, g0, false);
  cur_bf16_ = 0;
  qr(qu);
  for (int it = 0; it < plan_.opt.power_iterations; ++it) {
    g1.clear();
    g2.clear();
    for (auto& J : jobs) {
      dev::GemmJob a;
      a.A = J.X;
      a.B = J.U;
      a.C = J.Z;
      a.m = J.rows;
      a.n = J.r;
      a.k = J.cols;
      a.lda = J.ldx;
      a.ldb = J.cols;
      a.ldc = J.rows;
      a.alpha = 1.0;
      a.beta = 0.0;
      g1.push_back(a);
      dev::GemmJob b;
      b.A = J.X;
      b.B = J.Z;
      b.C = J.U;
      b.m = J.cols;
      b.n = J.r;
      b.k = J.rows;
      b.lda = J.ldx;
      b.ldb = J.rows;
      b.ldc = J.cols;
      b.alpha = 1.0;
      b.beta = 0.0;
      g2.push_back(b);
    }
    cur_bf16_ = bf16_sketch() ? 1 : 0;
    emit_gemm(OP_GEMM_NN, g1, false);
    cur_bf16_ = 0;
    qr(qz);
    cur_bf16_ = bf16_sketch() ? 1 : 0;
    emit_gemm(OP_GEMM_TN, g2, false);
    cur_bf16_ = 0;
    qr(qu);
  }
  for (auto& J : jobs) {
    dev::GemmJob g;
    g.A = J.X;
    g.B = J.U;
    g.C = J.V;
    g.m = J.rows;
    g.n = J.r;
    g.k = J.cols;
    g.lda = J.ldx;
    g.ldb = J.cols;
    g.ldc = J.rows;
    g.alpha = 1.0;
    g.beta = 0.0;
    g3.push_back(g);
  }
  emit_gemm(OP_GEMM_NN, g3, false);
}

// C(rows x rows, lower) -= V (U^T U) V^T
template <typename T>
void Engine<T>::lowrank_update(char* C, int ldc, int rows, const char* V,
                               const char* U, int cols, int r, char* K, char* Tm) {
  dev::GemmJob g;
  g.A = U;
  g.B = U;
  g.C = K;
  g.m = r;
  g.n = r;
  g.k = cols;
  g.lda = g.ldb = cols;
  g.ldc = r;
  g.alpha = 1.0;
  g.beta = 0.0;
  emit_gemm(OP_GEMM_TN, {g}, false);
  g.A = V;
  g.B = K;
  g.C = Tm;
  g.m = rows;
  g.n = r;
  g.k = r;
  g.lda = rows;
  g.ldb = r;
  g.ldc = rows;
  emit_gemm(OP_GEMM_NN, {g}, false);
  g.A = Tm;
  g.B = V;
  g.C = C;
  g.m = rows;
  g.n = rows;
  g.k = r;
  g.lda = rows;
  g.ldb = rows;
  g.ldc = ldc;
  g.alpha = -1.0;
  g.beta = 1.0;
  emit_gemm(OP_GEMM_NT, {g}, true);
}

// X (m x width of node q) <- X L_q^{-T} through the HODLR tree
// (DiagonalBlock.cpp:1151-1300, right-side transpose branch).
template <typename T>
void Engine<T>::hsolve_right(const Supernode& s, int q, char* X, int m, int ldx,
                             char* Tm) {
  const HNode& h = s.hnodes[q];
  if (h.leaf()) {
    int b = h.r1 - h.r0;
    trsm_right({{X, m, ldx, fp(h.off_leaf), b, b}});
    return;
  }
  const HNode& L = s.hnodes[h.left];
  const HNode& R = s.hnodes[h.right];
  const int cl = L.r1 - L.r0, cr = R.r1 - R.r0;
  char* Xl = (char*)((T*)X + (int64_t)(L.r0 - h.r0) * ldx);
  char* Xr = (char*)((T*)X + (int64_t)(R.r0 - h.r0) * ldx);
  hsolve_right(s, h.left, Xl, m, ldx, Tm);
  dev::GemmJob g;
  g.A = Xl;
  g.B = fp(h.off_U);
  g.C = Tm;
  g.m = m;
  g.n = h.rank;
  g.k = cl;
  g.lda = ldx;
  g.ldb = cl;
  g.ldc = m;
  g.alpha = 1.0;
  g.beta = 0.0;
  emit_gemm(OP_GEMM_NN, {g}, false);
  g.A = Tm;
  g.B = fp(h.off_V);
  g.C = Xr;
  g.m = m;
  g.n = cr;
  g.k = h.rank;
  g.lda = m;
  g.ldb = cr;
  g.ldc = ldx;
  g.alpha = -1.0;
  g.beta = 1.0;
  emit_gemm(OP_GEMM_NT, {g}, false);
  hsolve_right(s, h.right, Xr, m, ldx, Tm);
}

template <typename T>
void Engine<T>::layout_scratch() {
  // per compressed supernode: Z (R r), W (r r), K (r r), Tm (max(R,c) rmax);
  // per internal tree node: Z (rows r), W, K (r r), Tm (rows r)
  int64_t off = 0;
  node_scratch_.assign(plan_.nodes.size(), -1);
  wx_scratch_.assign(plan_.nodes.size(), 0);
  wx_elems_.assign(plan_.nodes.size(), 0);
  hnode_scratch_.assign(plan_.nodes.size(), {});
  inv_elems_ = 0;
  for (auto& lv : plan_.levels) {
    int64_t unc = 0;
    for (int j : lv)
      if (!plan_.nodes[j].compress) unc += (int64_t)plan_.nodes[j].ncols * plan_.nodes[j].ncols;
    inv_elems_ = std::max(inv_elems_, unc);
  }
  // compressed supernodes of one level are emitted together, each with its
  // own leaf-inverse and FP64-QR scratch
  inv_need_.assign(plan_.nodes.size(), 0);
  qr64_need_.assign(plan_.nodes.size(), 0);
  for (size_t j = 0; j < plan_.nodes.size(); ++j) {
    const Supernode& s = plan_.nodes[j];
    if (!s.compress) continue;
    int64_t lv = 0;
    for (auto& h : s.hnodes)
      if (h.leaf()) lv += (int64_t)(h.r1 - h.r0) * (h.r1 - h.r0);
    inv_need_[j] = lv;
  }
  for (auto& lv : plan_.levels) {
    int64_t sum = 0;
    for (int j : lv) sum += inv_need_[j];
    inv_elems_ = std::max(inv_elems_, sum);
  }
  // per-supernode sketch, helper and tree scratch is used only while its
  // level is emitted: offsets restart at every level (the supernodes of one
  // level share launches, so theirs are disjoint)
  int64_t scratch_max = 0;
  for (auto& lvl : plan_.levels)
  for (int j : lvl) {
    if (j == lvl.front()) off = 0;
    const Supernode& s = plan_.nodes[j];
    if (!s.compress) continue;
    int rmax = s.rank;
    for (auto& h : s.hnodes) rmax = std::max(rmax, h.rank);
    node_scratch_[j] = off;
    const int64_t R = s.R(), r = s.rank;
    off += R * r + 2 * r * r + (int64_t)std::max<int64_t>(R, s.ncols) * rmax;
    hnode_scratch_[j].assign(s.hnodes.size(), -1);
    {
      int maxh = 0;
      for (auto& h : s.hnodes) maxh = std::max(maxh, h.height);
      const std::vector<HTarget> tg = helper_targets((int)j);
      int64_t best = 0;
      for (int hh = 1; hh <= maxh; ++hh) {
        int64_t sum = 0;
        for (const HTarget& t : tg)
          for (size_t q = 0; q < s.hnodes.size(); ++q) {
            const HNode& hn = s.hnodes[q];
            if (hn.leaf() || hn.height != hh) continue;
            if (inside_subtree(s, (int)q, t.sub)) sum += (int64_t)hn.rank * t.rank;
          }
        best = std::max(best, sum);
      }
      wx_scratch_[j] = off;
      wx_elems_[j] = best;
      off += best;
    }
    for (size_t q = 0; q < s.hnodes.size(); ++q) {
      const HNode& h = s.hnodes[q];
      if (h.leaf()) continue;
      const int64_t rows = s.hnodes[h.right].r1 - s.hnodes[h.right].r0;
      hnode_scratch_[j][q] = off;
      off += 2 * rows * h.rank + 2 * (int64_t)h.rank * h.rank;
    }
    scratch_max = std::max(scratch_max, off);
  }
  scratch_elems_ = scratch_max;
  // Zp staging: per level, the uncompressed panels' R x c blocks
  zp_elems_ = 0;
  for (auto& lv : plan_.levels) {
    int64_t sum = 0;
    for (int j : lv)
      if (!plan_.nodes[j].compress) sum += (int64_t)plan_.nodes[j].R() * plan_.nodes[j].ncols;
    zp_elems_ = std::max(zp_elems_, sum);
  }
  // FP64 QR scratch (FP32 factors): largest concurrent batch of QRs, each
  // max(rows, cols) x r + r x r doubles, counted in T elements
  qr64_elems_ = 0;
  if (sizeof(T) == 4) {
    for (size_t j = 0; j < plan_.nodes.size(); ++j) {
      const Supernode& s = plan_.nodes[j];
      if (!s.compress) continue;
      int64_t mx = 0;
      for (const HStep& st : s.hsteps) {
        int64_t sum = 0;
        std::function<void(int)> acc = [&](int q) {
          const HNode& h = s.hnodes[q];
          if (h.leaf()) return;
          int64_t rows = s.hnodes[h.right].r1 - s.hnodes[h.right].r0;
          int64_t cols = s.hnodes[h.left].r1 - s.hnodes[h.left].r0;
          sum += std::max(rows, cols) * h.rank + (int64_t)h.rank * h.rank;
          if (st.kind == HStep::EXPLICIT) {
            acc(h.left);
            acc(h.right);
          }
        };
        if (st.kind != HStep::LEAF) acc(st.node);
        mx = std::max(mx, sum);
      }
      mx = std::max(mx, (int64_t)std::max(s.R(), s.ncols) * s.rank + (int64_t)s.rank * s.rank);
      qr64_need_[j] = 2 * mx;
    }
    for (auto& lv : plan_.levels) {
      int64_t sum = 0;
      for (int j : lv) sum += qr64_need_[j];
      qr64_elems_ = std::max(qr64_elems_, sum);
    }
  }
}

// The compressed supernodes of one level are independent (each writes its
// own panel, factor storage, update matrix and scratch), so their step lists
// are recorded separately and merged: each merged launch takes the next step
// of every supernode whose next step has the same kind, choosing first the
// kind at the head of the longest remaining list.  Each supernode's own
// order is kept.  RSC_FACTOR_BATCH=0 emits them one after another instead.
template <typename T>
void Engine<T>::emit_level_compressed(const std::vector<int>& nodes) {
  const char* e = std::getenv("RSC_FACTOR_BATCH");
  const bool batch = !(e && e[0] == '0') && nodes.size() > 1;
  int64_t inv_off = inv_base(), q64_off = qr64_base();
  if (!batch) {
    for (int j : nodes) {
      inv_cursor_ = inv_base();
      qr64_cursor_ = qr64_base();
      emit_compressed(j);
    }
    return;
  }
  std::vector<std::vector<Pending>> q(nodes.size());
  for (size_t i = 0; i < nodes.size(); ++i) {
    inv_cursor_ = inv_off;
    qr64_cursor_ = q64_off;
    inv_off += inv_need_[nodes[i]];
    q64_off += qr64_need_[nodes[i]];
    capture_ = &q[i];
    try {
      emit_compressed(nodes[i]);
    } catch (...) {
      capture_ = nullptr;
      throw;
    }
    capture_ = nullptr;
  }
  if (cur_f64_ != 0) throw std::logic_error("unbalanced FP64 section");
  std::vector<size_t> head(nodes.size(), 0);
  auto same = [](const Pending& a, const Pending& b) {
    return a.kind == b.kind && a.lower == b.lower && a.f64 == b.f64 && a.bf16 == b.bf16;
  };
  for (;;) {
    size_t lead = nodes.size(), most = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
      const size_t left = q[i].size() - head[i];
      if (left > most) {
        most = left;
        lead = i;
      }
    }
    if (lead == nodes.size()) break;
    Pending m = q[lead][head[lead]];
    m.g.clear(); m.p.clear(); m.t.clear(); m.c.clear(); m.e.clear(); m.v.clear(); m.s.clear();
    int merged = 0;
    for (size_t i = 0; i < nodes.size(); ++i) {
      if (head[i] >= q[i].size()) continue;
      const Pending& h = q[i][head[i]];
      if (!same(h, m)) continue;
      m.g.insert(m.g.end(), h.g.begin(), h.g.end());
      m.p.insert(m.p.end(), h.p.begin(), h.p.end());
      m.t.insert(m.t.end(), h.t.begin(), h.t.end());
      m.c.insert(m.c.end(), h.c.begin(), h.c.end());
      m.e.insert(m.e.end(), h.e.begin(), h.e.end());
      m.v.insert(m.v.end(), h.v.begin(), h.v.end());
      m.s.insert(m.s.end(), h.s.begin(), h.s.end());
      ++head[i];
      ++merged;
    }
    cur_f64_ = m.f64;
    cur_bf16_ = m.bf16;
    ctx_ = m.ctx + (merged > 1 ? " (+" + std::to_string(merged - 1) + " supernodes)" : "");
    switch (m.kind) {
      case OP_GEMM_NN:
      case OP_GEMM_TN:
      case OP_GEMM_NT:
      case OP_GEMM_TT: emit_gemm(m.kind, m.g, m.lower); break;
      case OP_POTRF: emit_potrf(m.p); break;
      case OP_TRSM: emit_trsm(m.t); break;
      case OP_COPY: emit_copy(m.c); break;
      case OP_EYE: emit_eye(m.e); break;
      case OP_CONV: emit_conv(m.v); break;
      case OP_SHIFT: emit_shift(m.s); break;
      default: throw std::logic_error("unexpected recorded step");
    }
  }
  cur_f64_ = 0;
  cur_bf16_ = 0;
}

template <typename T>
void Engine<T>::emit_compressed(int j) {
  const Supernode& s = plan_.nodes[j];
  const int c = s.ncols, R = s.R(), ld = c + R;
  T* P = Wb() + s.off_work_panel;
  const int64_t nsb = scratch_base() + node_scratch_[j];
  int rmax = s.rank;
  for (auto& h : s.hnodes) rmax = std::max(rmax, h.rank);
  char* Tm = wp(nsb + (int64_t)R * s.rank + 2 * (int64_t)s.rank * s.rank);
  if (rng_step_ < 0) rng_step_ = (int)prog_.size();
  const T* G = Wb() + rng_base();
  auto hsk = [&](int q) -> SketchJob {
    const HNode& h = s.hnodes[q];
    const HNode& L = s.hnodes[h.left];
    const HNode& Rt = s.hnodes[h.right];
    const int rows = Rt.r1 - Rt.r0, cols = L.r1 - L.r0;
    const int64_t sb = scratch_base() + hnode_scratch_[j][q];
    SketchJob J;
    J.X = (char*)(P + Rt.r0 + (int64_t)L.r0 * ld);
    J.rows = rows;
    J.cols = cols;
    J.ldx = ld;
    J.G = (const char*)(G + h.rng_off);
    J.r = h.rank;
    J.U = fp(h.off_U);
    J.V = fp(h.off_V);
    J.Z = wp(sb);
    J.W = wp(sb + (int64_t)rows * h.rank);
    J.node = j;
    return J;
  };
  for (const HStep& st : s.hsteps) {
    const HNode& h = s.hnodes[st.node];
    const int b = h.r1 - h.r0;
    ctx_ = "node " + std::to_string(j) + " diag step kind " + std::to_string((int)st.kind) +
           " tree node " + std::to_string(st.node);
    if (st.kind == HStep::EXPLICIT || st.kind == HStep::LEAF) {
      panel_chol({{(char*)(P + h.r0 + (int64_t)h.r0 * ld), b, b, ld, j, factor_tol()}});
      std::vector<SketchJob> sk;
      std::vector<dev::CopyJob> cp;
      std::function<void(int)> pre = [&](int q) {
        const HNode& x = s.hnodes[q];
        if (x.leaf()) {
          int bb = x.r1 - x.r0;
          cp.push_back({P + x.r0 + (int64_t)x.r0 * ld, F() + x.off_leaf, bb, bb, ld, bb});
          return;
        }
        sk.push_back(hsk(q));
        pre(x.left);
        pre(x.right);
      };
      pre(st.node);
      if (!sk.empty()) sketch(sk);
      emit_copy(cp);
    } else {
      // recursive off-diagonal block: L_N = S_N L_left^{-T}, compress, update
      SketchJob J = hsk(st.node);
      hsolve_right(s, h.left, J.X, J.rows, ld, Tm);
      sketch({J});
      const HNode& Rt = s.hnodes[h.right];
      const int64_t sb = scratch_base() + hnode_scratch_[j][st.node];
      char* K = wp(sb + (int64_t)J.rows * h.rank + (int64_t)h.rank * h.rank);
      char* T2 = wp(sb + (int64_t)J.rows * h.rank + 2 * (int64_t)h.rank * h.rank);
      lowrank_update((char*)(P + Rt.r0 + (int64_t)Rt.r0 * ld), ld, J.rows, J.V, J.U,
                     J.cols, h.rank, K, T2);
    }
  }
  if (R > 0) {
    int root = 0;
    for (int q = 0; q < (int)s.hnodes.size(); ++q)
      if (s.hnodes[q].parent < 0) root = q;
    char* X = (char*)(P + c);
    ctx_ = "node " + std::to_string(j) + " L21 solve";
    hsolve_right(s, root, X, R, ld, Tm);
    ctx_ = "node " + std::to_string(j) + " off-diagonal sketch";
    SketchJob J;
    J.X = X;
    J.rows = R;
    J.cols = c;
    J.ldx = ld;
    J.G = (const char*)(G + s.rng_off);
    J.r = s.rank;
    J.U = fp(s.off_U);
    J.V = fp(s.off_V);
    J.Z = wp(nsb);
    J.W = wp(nsb + (int64_t)R * s.rank);
    J.node = j;
    sketch({J});
    char* K = wp(nsb + (int64_t)R * s.rank + (int64_t)s.rank * s.rank);
    lowrank_update(wp(upd_base() + s.off_update), R, R, J.V, J.U, c, s.rank, K, Tm);
  }
  if (s.off_dinv >= 0) {
    // dense L_D^{-T} = I L_D^{-T} through the tree (leaves still hold L)
    ctx_ = "node " + std::to_string(j) + " dense diagonal inverse";
    emit_eye({{F() + s.off_dinv, c, c}});
    hsolve_right(s, hodlr_root(s), fp(s.off_dinv), c, c, Tm);
  }
  ctx_ = "node " + std::to_string(j) + " invert leaves";
  std::vector<CholMat> leaves;
  for (auto& h : s.hnodes)
    if (h.leaf()) leaves.push_back({fp(h.off_leaf), h.r1 - h.r0, h.r1 - h.r0, h.r1 - h.r0, j, 0});
  invert(leaves);
  ctx_ = "node " + std::to_string(j) + " solve helpers";
  emit_solve_helpers(j);
}

template <typename T>
std::vector<typename Engine<T>::HTarget> Engine<T>::helper_targets(int j) const {
  const Supernode& s = plan_.nodes[j];
  std::vector<HTarget> tg;
  for (int M = 0; M < (int)s.hnodes.size(); ++M) {
    const HNode& hm = s.hnodes[M];
    if (hm.leaf()) continue;
    tg.push_back({hm.right, true, hm.off_V, hm.off_W, hm.rank});
    tg.push_back({hm.left, false, hm.off_U, hm.off_X, hm.rank});
  }
  if (s.R() > 0 && s.rank > 0) tg.push_back({hodlr_root(s), false, s.off_U, s.off_Xoff, s.rank});
  return tg;
}

// Solve helpers of a compressed diagonal.  For every internal tree node M
// with children (A, B): W_M = L_B^{-1} V_M and X_M = L_A^{-T} U_M; for the
// off-diagonal pair X_off = L_D^{-T} U.  With them the tree solves become
// one leaf step plus one low-rank correction per tree height (forward
// y_B -= W_M (U_M^T y_A), transposed x_A -= X_M (V_M^T x_B), bottom-up),
// and the off-diagonal block acts on the front without a diagonal solve:
// L21 L_D^{-1} = V X_off^T.  Computed with the same recursions applied to
// the target matrices: leaf blocks first (leaves hold L^{-T}), then the
// corrections of every internal node N inside the target's subtree, by
// height.
template <typename T>
void Engine<T>::emit_solve_helpers(int j) {
  const Supernode& s = plan_.nodes[j];
  int maxh = 0;
  for (auto& h : s.hnodes) maxh = std::max(maxh, h.height);
  const std::vector<HTarget> tg = helper_targets(j);
  if (tg.empty()) return;
  auto sub_r0 = [&](const HTarget& t) { return s.hnodes[t.sub].r0; };
  auto sub_n = [&](const HTarget& t) { return s.hnodes[t.sub].r1 - s.hnodes[t.sub].r0; };
  std::vector<dev::GemmJob> tn, nn;
  for (const HTarget& t : tg) {
    const int ld = sub_n(t), r0 = sub_r0(t);
    for (int q = 0; q < (int)s.hnodes.size(); ++q) {
      const HNode& lf = s.hnodes[q];
      if (!lf.leaf() || !inside_subtree(s, q, t.sub)) continue;
      const int b = lf.r1 - lf.r0;
      dev::GemmJob g;
      g.m = b;
      g.n = t.rank;
      g.k = b;
      g.lda = b;
      g.alpha = 1.0;
      g.beta = 0.0;
      g.A = fp(lf.off_leaf);
      g.B = fp(t.src + (lf.r0 - r0));
      g.C = fp(t.dst + (lf.r0 - r0));
      g.ldb = g.ldc = ld;
      (t.isW ? tn : nn).push_back(g);
    }
  }
  emit_gemm(OP_GEMM_TN, tn, false);
  emit_gemm(OP_GEMM_NN, nn, false);
  const int64_t tbase = scratch_base() + wx_scratch_[j];
  for (int h = 1; h <= maxh; ++h) {
    std::vector<dev::GemmJob> t1, t2, u1, u2;
    int64_t toff = tbase;
    for (const HTarget& t : tg) {
      const int ld = sub_n(t), r0 = sub_r0(t), rm = t.rank;
      for (int N = 0; N < (int)s.hnodes.size(); ++N) {
        const HNode& hn = s.hnodes[N];
        if (hn.leaf() || hn.height != h || !inside_subtree(s, N, t.sub)) continue;
        const HNode& NA = s.hnodes[hn.left];
        const HNode& NB = s.hnodes[hn.right];
        const int ncl = NA.r1 - NA.r0, ncr = NB.r1 - NB.r0, rn = hn.rank;
        char* Tm = wp(toff);
        toff += (int64_t)rn * rm;
        dev::GemmJob g;
        g.alpha = 1.0;
        g.beta = 0.0;
        dev::GemmJob u;
        u.B = Tm;
        u.n = rm;
        u.k = rn;
        u.ldb = rn;
        u.ldc = ld;
        u.alpha = -1.0;
        u.beta = 1.0;
        if (t.isW) {
          // T = U_N^T D[N.A rows]; D[N.B rows] -= W_N T
          g.A = fp(hn.off_U);
          g.B = fp(t.dst + (NA.r0 - r0));
          g.m = rn;
          g.n = rm;
          g.k = ncl;
          g.lda = ncl;
          g.ldb = ld;
          g.C = Tm;
          g.ldc = rn;
          t1.push_back(g);
          u.A = fp(hn.off_W);
          u.C = fp(t.dst + (NB.r0 - r0));
          u.m = ncr;
          u.lda = ncr;
          u1.push_back(u);
        } else {
          // T = V_N^T D[N.B rows]; D[N.A rows] -= X_N T
          g.A = fp(hn.off_V);
          g.B = fp(t.dst + (NB.r0 - r0));
          g.m = rn;
          g.n = rm;
          g.k = ncr;
          g.lda = ncr;
          g.ldb = ld;
          g.C = Tm;
          g.ldc = rn;
          t2.push_back(g);
          u.A = fp(hn.off_X);
          u.C = fp(t.dst + (NA.r0 - r0));
          u.m = ncl;
          u.lda = ncl;
          u2.push_back(u);
        }
      }
    }
    if (toff - tbase > wx_elems_[j]) throw std::runtime_error("helper scratch too small");
    emit_gemm(OP_GEMM_TN, t1, false);
    emit_gemm(OP_GEMM_TN, t2, false);
    emit_gemm(OP_GEMM_NN, u1, false);
    emit_gemm(OP_GEMM_NN, u2, false);
  }
}

// Timing trace of one numeric factorization: one line per program step
// (index, operation, workgroups launched, floating-point operations of the
// step's jobs, milliseconds between the events around it, context).
template <typename T>
void Engine<T>::write_factor_trace(const char* path, std::vector<hipEvent_t>& ev) {
  static const char* names[] = {"assemble", "gemm_nn", "gemm_tn", "gemm_nt", "gemm_tt",
                                "potrf", "trsm", "copy", "shift", "eye", "check", "conv"};
  FILE* fp = std::fopen(path, "w");
  for (size_t si = 0; fp && si < prog_.size(); ++si) {
    const Step& st = prog_[si];
    float ms = 0.f;
    RSC_HIP(hipEventElapsedTime(&ms, ev[si], ev[si + 1]));
    double flops = 0;
    int grid = st.count;
    if (st.op >= OP_GEMM_NN && st.op <= OP_GEMM_TT) {
      for (int t = 0; t < st.count; ++t) {
        const dev::GemmTile& tl = gtiles_[st.first + t];
        const dev::GemmJob& J = gjobs_[tl.job];
        const double r = std::min(dev::TILE, J.m - tl.tm * dev::TILE);
        const double c = std::min(dev::TILE, J.n - tl.tn * dev::TILE);
        flops += 2.0 * r * c * (tl.k1 - tl.k0);
      }
    } else if (st.op == OP_POTRF) {
      for (int t = 0; t < st.count; ++t) {
        const double n = pjobs_[st.first + t].n;
        flops += n * n * n / 3.0;
      }
    } else if (st.op == OP_TRSM) {
      for (int t = 0; t < st.count; ++t) {
        const dev::RowTile& tl = ttiles_[st.first + t];
        const dev::TrsmJob& J = tjobs_[tl.job];
        const double rows = std::min(dev::TILE, J.m - tl.row0);
        flops += rows * J.nb * J.nb;
      }
    }
    std::fprintf(fp, "%zu %s%s%s %d %.0f %.4f %s\n", si, names[st.op], st.f64 ? "64" : "",
                 st.bf16 ? "bf" : "", grid, flops, ms, step_ctx_[si].c_str());
  }
  if (fp) std::fclose(fp);
  for (auto& e : ev) RSC_HIP(hipEventDestroy(e));
  ev.clear();
}

// Debug: after every program step, scan that step's outputs for non-finite
// values; the first failing step index is reported with its description.
template <typename T>
void Engine<T>::add_checks() {
  std::vector<Step> out;
  step_desc_.clear();
  const char* names[] = {"assemble", "gemm_nn", "gemm_tn", "gemm_nt", "gemm_tt",
                         "potrf", "trsm", "copy", "shift", "eye", "check", "conv"};
  for (size_t si = 0; si < prog_.size(); ++si) {
    const Step& st = prog_[si];
    out.push_back(st);
    const int tag = (int)step_desc_.size();
    std::vector<dev::CheckJob> jobs;
    std::string desc = names[st.op];
    if (st.op >= OP_GEMM_NN && st.op <= OP_GEMM_TT) {
      int last = -1;
      for (int t = 0; t < st.count; ++t) {
        const dev::GemmTile& tl = gtiles_[st.first + t];
        if (tl.job == last) continue;
        last = tl.job;
        const dev::GemmJob& J = gjobs_[tl.job];
        jobs.push_back({J.C, J.m, J.n, J.ldc, tag});
      }
    } else if (st.op == OP_POTRF) {
      for (int t = 0; t < st.count; ++t) {
        const dev::PotrfJob& J = pjobs_[st.first + t];
        jobs.push_back({J.A, J.n, J.n, J.lda, tag});
        if (t == 0) desc += " node " + std::to_string(J.node);
      }
    } else if (st.op == OP_TRSM) {
      int last = -1;
      for (int t = 0; t < st.count; ++t) {
        const dev::RowTile& tl = ttiles_[st.first + t];
        if (tl.job == last) continue;
        last = tl.job;
        const dev::TrsmJob& J = tjobs_[tl.job];
        jobs.push_back({J.B, J.m, J.nb, J.ldb, tag});
      }
    } else if (st.op == OP_COPY) {
      int last = -1;
      for (int t = 0; t < st.count; ++t) {
        const dev::RowTile& tl = ctiles_[st.first + t];
        if (tl.job == last) continue;
        last = tl.job;
        const dev::CopyJob& J = cjobs_[tl.job];
        jobs.push_back({J.dst, J.m, J.n, J.ldd, tag});
      }
    } else if (st.op == OP_SHIFT) {
      for (int t = 0; t < st.count; ++t) {
        const dev::ShiftJob& J = sjobs_[st.first + t];
        jobs.push_back({J.W, J.r, J.r, J.ld, tag});
      }
    }
    desc += " (step " + std::to_string(si) + ", " + std::to_string(jobs.size()) +
            " jobs) in " + step_ctx_[si];
    if (!jobs.empty())
      desc += " first job " + std::to_string(jobs[0].rows) + "x" +
              std::to_string(jobs[0].cols) + " ld " + std::to_string(jobs[0].ld);
    step_desc_.push_back(desc);
    if (jobs.empty()) continue;
    Step ck{OP_CHECK, (int)jobs.size(), (int64_t)kjobs_.size(), st.f64};
    kjobs_.insert(kjobs_.end(), jobs.begin(), jobs.end());
    out.push_back(ck);
  }
  prog_.swap(out);
}

// Joins the dataflow task builder and uploads its lists (once).
template <typename T>
void Engine<T>::finish_solve_setup() {
  if (solve_ready_) return;
  if (df_thread_.joinable()) df_thread_.join();
  if (df_err_) {
    std::exception_ptr e = df_err_;
    df_err_ = nullptr;
    std::rethrow_exception(e);
  }
  const char* tp = std::getenv("RSC_DF_TRACE");
  if (tp && tp[0]) df_trace_path_ = tp;
  int64_t slots = 0, maxtasks = 0;
  for (int q = 0; q < 2; ++q) {
    DataflowTasks& DT = q == 0 ? *df_pending_ : *df_pending2_;
    DfList& D = df_[q];
    D.ntasks = (int)DT.tasks.size();
    D.ngroups = (int)DT.need.size();
    D.tasks.upload(DT.tasks, st_);
    D.gv.upload(DT.gv, st_);
    std::vector<dev::DfWait> waits(DT.gw.size());
    for (size_t q = 0; q < DT.gw.size(); ++q) waits[q] = {DT.gw[q], DT.need[DT.gw[q]]};
    D.gw.upload(waits, st_);
    D.cnt.alloc((size_t)(D.ngroups + 1) * sizeof(unsigned));
    if (df_queue_) {
      D.q_queue.alloc((size_t)std::max(1, D.ntasks) * sizeof(int));
      D.q_ctr.alloc(2 * sizeof(unsigned));
      D.q_pend.alloc((size_t)std::max(1, D.ngroups) * sizeof(int));
      D.q_pend0.upload(DT.q_pend0, st_);
      D.q_need.upload(DT.need, st_);
      D.q_gptr.upload(DT.q_gptr, st_);
      D.q_gtasks.upload(DT.q_gtasks, st_);
      D.q_ginfo.upload(DT.q_ginfo.empty() ? std::vector<int>(4, 0) : DT.q_ginfo, st_);
      D.q_sptr.upload(DT.q_sptr, st_);
      D.q_succ.upload(DT.q_succ.empty() ? std::vector<int>{0} : DT.q_succ, st_);
      D.q_init.upload(DT.q_init.empty() ? std::vector<int>{0} : DT.q_init, st_);
      D.q_n0 = (int)DT.q_init.size();
    }
    D.grid = std::max(1, std::min(D.ntasks, df_resident_[q]));
    if (!df_trace_path_.empty()) {
      D.host = DT.tasks;
      D.host_wait = waits;
      D.host_gv = DT.gv;
      D.host_feat = DT.trace_feat;
    }
    slots = std::max(slots, DT.slots);
    maxtasks = std::max<int64_t>(maxtasks, D.ntasks);
  }
  if (!df_trace_path_.empty())
    dtrace_.alloc((size_t)4 * maxtasks * sizeof(unsigned long long));
  part_stride_ = std::max<int64_t>(slots * 32, 1);
  df_pending_.reset();
  df_pending2_.reset();
  RSC_HIP(hipStreamSynchronize(st_));
  solve_ready_ = true;
}

template <typename T>
void Engine<T>::setup_static() {
  const Plan& P = plan_;
  const int ns = (int)P.nodes.size();
  auto tp = std::chrono::steady_clock::now();
  auto phase = [&](const char* what) {
    if (!P.opt.verbose) return;
    auto now = std::chrono::steady_clock::now();
    std::printf("setup %s: %.3f s\n", what, std::chrono::duration<double>(now - tp).count());
    tp = now;
  };
  // the factor arena and the scratch layout first: the numeric program is
  // then built on a host thread while the rest of the setup runs here
  {
    const char* dbg = std::getenv("RSC_DEBUG_CHECK");
    debug_check_ = dbg && dbg[0] == '1';
  }
  factor_.alloc(std::max<int64_t>(P.factor_elems, 1) * sizeof(T));
  layout_scratch();
  phase("factor allocation and scratch layout");
  prog_thread_ = std::thread([this]() {
    try {
      build_program();
    } catch (...) {
      prog_err_ = std::current_exception();
    }
  });
  // permuted lower pattern and the destination of every entry in its panel
  LowerCsc L = permuted_lower(A_, P.pinv);
  std::vector<int64_t> a_dst(L.i.size());
  {
    // front position of a row, one supernode at a time per thread
    std::vector<std::vector<int>> rposes(stream_threads());
    parallel_for(ns, 256, [&](int64_t b, int64_t e, int t) {
      std::vector<int>& rpos = rposes[t];
      if (rpos.empty()) rpos.assign(P.n, -1);
      for (int64_t j = b; j < e; ++j) {
        const Supernode& s = P.nodes[j];
        const int c = s.ncols, ld = c + s.R(), last = s.first + c - 1;
        for (size_t q = 0; q < s.rows.size(); ++q) rpos[s.rows[q]] = c + (int)q;
        for (int k = s.first; k <= last; ++k)
          for (int x = L.p[k]; x < L.p[k + 1]; ++x) {
            const int r = L.i[x];
            const int pos = r <= last ? r - s.first : rpos[r];
            a_dst[x] = (int64_t)pos + (int64_t)(k - s.first) * ld;
          }
        for (int r : s.rows) rpos[r] = -1;
      }
    });
  }
  dlowp_.upload(L.p, st_);
  da_dst_.upload(a_dst, st_);
  da_src_.upload(L.src, st_);
  phase("permuted lower matrix");
  // solve: vector layout, level-scheduled and dataflow task lists
  SolveLayout& SL = sl_;
  {
    // subtree tasks: up to 16384 panel elements each (measured best for one
    // and for sixteen right-hand sides at 64^3 and 128^3)
    int64_t lim = 16384;
    if (const char* e = std::getenv("RSC_SUBTREE_ELEMS")) lim = std::atoll(e);
    // multi-right-hand-side subtree tasks: the supernode limits (front and
    // y in LDS) bound them; RSC_MS_SUBTREE_ELEMS caps their panel elements
    int64_t mlim = 1 << 20;
    if (const char* e = std::getenv("RSC_MS_SUBTREE_ELEMS")) mlim = std::atoll(e);
    // at most as many multi-right-hand-side subtree tasks as the dataflow
    // launch has resident workgroups (the smallest are merged), so all of
    // them start at once; RSC_MS_SUBTREE_TASKS overrides (0: no merging)
    int mtasks = 0;
    {
      int dev = 0, cus = 1;
      RSC_HIP(hipGetDevice(&dev));
      RSC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      int cap = 4;
      if (const char* e = std::getenv("RSC_DF_BLOCKS_PER_CU")) cap = std::max(1, std::atoi(e));
      mtasks = cus * std::min(cap, dev::dataflow_blocks_per_cu<T>(true));
    }
    if (const char* e = std::getenv("RSC_MS_SUBTREE_TASKS")) mtasks = std::atoi(e);
    build_solve_layout(P, SL, lim, mlim, mtasks);
  }
  phase("solve layout");
  nsub_ = (int)SL.sub.root.size();
  dstptr_.upload(SL.sub.ptr, st_);
  dstnodes_.upload(SL.sub.nodes.empty() ? std::vector<int>{0} : SL.sub.nodes, st_);
  dstcols_.upload(SL.sub.cols.empty() ? std::vector<int>{0, -1} : SL.sub.cols, st_);
  dmsptr_.upload(SL.sub_multi.ptr, st_);
  {
    // the multi-right-hand-side subtree tasks read a task root as ~j (its
    // update vector is stored for other workgroups)
    std::vector<int> mn(SL.sub_multi.nodes);
    for (int& j : mn) {
      const int par = P.nodes[j].parent;
      if (par < 0 || SL.sub_multi.of[par] != SL.sub_multi.of[j]) j = ~j;
    }
    dmsnodes_.upload(mn.empty() ? std::vector<int>{0} : mn, st_);
  }
  dmscols_.upload(SL.sub_multi.cols.empty() ? std::vector<int>{0, -1} : SL.sub_multi.cols, st_);
  dpanel_.upload(SL.panel, st_);
  dchildren_.upload(SL.children, st_);
  drelmap_.upload(SL.relmap, st_);
  dranges_.upload(SL.asm_ranges, st_);
  drows_.upload(SL.rows, st_);
  std::vector<int> permv(P.perm);
  dperm_.upload(permv, st_);
  std::vector<int> pinvv(P.pinv);
  dpinv_.upload(pinvv, st_);
  fv_off_ = SL.fv_off;
  fv_stride_ = SL.fv_stride;
  tv_stride_ = SL.tv_stride;
  dvnodes_.upload(SL.vn, st_);
  {
    // one persistent dataflow launch per apply; RSC_APPLY_LEVELS=1 selects the
    // level-scheduled launches instead
    const char* lv = std::getenv("RSC_APPLY_LEVELS");
    use_dataflow_ = !(lv && lv[0] == '1');
    // workgroups take ready tasks from a queue; RSC_DF_QUEUE=0 hands out
    // tickets in list-schedule order instead
    const char* qv = std::getenv("RSC_DF_QUEUE");
    df_queue_ = !(qv && qv[0] == '0');
  }
  if (use_dataflow_) {
    int dev = 0, cus = 1;
    RSC_HIP(hipGetDevice(&dev));
    RSC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // RSC_DF_BLOCKS_PER_CU caps the workgroups per compute unit (default 4:
    // more co-resident workgroups lengthen every task on the critical path)
    int cap = 4;
    if (const char* e = std::getenv("RSC_DF_BLOCKS_PER_CU")) cap = std::max(1, std::atoi(e));
    df_resident_[0] = cus * std::min(cap, dev::dataflow_blocks_per_cu<T>(false));
    df_resident_[1] = cus * std::min(cap, dev::dataflow_blocks_per_cu<T>(true));
    // ticket order: a list schedule on the resident workgroups of each kernel
    // (costs for one and for sixteen right-hand sides); RSC_TICKET_ORDER=0
    // keeps level order
    TicketModel one, many;
    const char* to = std::getenv("RSC_TICKET_ORDER");
    const bool level_order = to && to[0] == '0';
    one.workers = level_order ? 0 : df_resident_[0];
    many.workers = level_order ? 0 : df_resident_[1];
    one.elem_bytes = many.elem_bytes = (int)sizeof(T);
    one.nrhs = 1;
    many.nrhs = 16;
    one.chunk_group = kSingleChunkGroup;
    many.chunk_group = 1;
    if (const char* e = std::getenv("RSC_DF_COL_ROWS")) one.col_rows = many.col_rows = std::atoi(e);
    if (const char* e = std::getenv("RSC_DF_ROW_COLS")) one.row_cols = many.row_cols = std::atoi(e);
    // the dataflow task lists are built on host threads while the device
    // factors; finish_solve_setup() joins them before the first apply
    df_thread_ = std::thread([this, one, many]() {
      std::exception_ptr err2;
      std::thread second([&]() {
        try {
          df_pending2_.reset(new DataflowTasks());
          build_dataflow_tasks(plan_, sl_, *df_pending2_, &sl_.sub_multi, many);
        } catch (...) {
          err2 = std::current_exception();
        }
      });
      try {
        df_pending_.reset(new DataflowTasks());
        build_dataflow_tasks(plan_, sl_, *df_pending_, &sl_.sub, one);
      } catch (...) {
        df_err_ = std::current_exception();
      }
      second.join();
      if (!df_err_ && err2) df_err_ = err2;
    });
  } else {
    LevelTasks LT;
    build_level_tasks(P, SL, LT);
    fwd_.swap(LT.fwd);
    bwd_.swap(LT.bwd);
    level_nodes_.swap(LT.level_nodes);
    dgv_.upload(LT.gv, st_);
    ntasks_ = (int64_t)LT.gv.size();
    part_stride_ = std::max<int64_t>(LT.slots * 32, 1);
    solve_ready_ = true;
  }
  phase("task lists");
  {
    int dev = 0, cus = 1, khz = 0;
    RSC_HIP(hipGetDevice(&dev));
    RSC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0)
      khz = 100000;
    df_timeout_ = (unsigned long long)khz * 1000ull * 4ull;  // 4 s
    // RSC_DF_TIMEOUT_US shortens the bound of every dependency wait (for
    // tests of the abort path: the launch drains and the apply reports
    // RSC_DEVICE_ERROR)
    // RSC_DF_BACKOFF: longest pause between polls of a dependency wait
    if (const char* e = std::getenv("RSC_DF_BACKOFF")) df_backoff_ = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("RSC_DF_TIMEOUT_US"))
      df_timeout_ = (unsigned long long)khz * (unsigned long long)std::atoll(e) / 1000ull;
    std::vector<int> z = {0, 0};
    dapst_.upload(z, st_);
  }
  std::vector<int> levnodes;
  lev_off_.clear();
  for (auto& lv : level_nodes_) {
    lev_off_.push_back((int)levnodes.size());
    levnodes.insert(levnodes.end(), lv.begin(), lv.end());
  }
  lev_off_.push_back((int)levnodes.size());
  lev_front_.clear();
  for (auto& lv : level_nodes_) {
    int mf = 1;
    for (int j : lv) mf = std::max(mf, P.nodes[j].ncols + P.nodes[j].R());
    lev_front_.push_back(mf);
  }
  if (levnodes.empty()) levnodes.push_back(0);
  dlevnodes_.upload(levnodes, st_);
  std::vector<int> status = {0, INT_MAX, 0, INT_MAX};
  dstatus_.upload(status, st_);
  RSC_HIP(hipStreamSynchronize(st_));
  phase("uploads");
}

template <typename T>
void Engine<T>::build_program() {
  const Plan& P = plan_;
  const int ns = (int)P.nodes.size();
  prog_.clear();
  step_ctx_.clear();
  gjobs_.clear();
  gtiles_.clear();
  ngcnt_ = 0;
  gpart_bytes_ = 0;
  pjobs_.clear();
  tjobs_.clear();
  ttiles_.clear();
  cjobs_.clear();
  ctiles_.clear();
  sjobs_.clear();
  ejobs_.clear();
  etiles_.clear();
  kjobs_.clear();
  vjobs_.clear();
  tasks_.clear();
  const int64_t welems = zp_base() + zp_elems_;
  if (P.opt.verbose)
    std::printf("workspace elements: work %lld update %lld scratch %lld rng %lld inverse %lld "
                "qr64 %lld zp %lld\n",
                (long long)P.work_elems, (long long)P.update_elems, (long long)scratch_elems_,
                (long long)P.rng_count, (long long)inv_elems_, (long long)qr64_elems_,
                (long long)zp_elems_);
  auto tw = std::chrono::steady_clock::now();
  work_.alloc(std::max<int64_t>(welems, 1) * sizeof(T));
  if (P.opt.verbose)
    std::printf("workspace allocation: %.3f s\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - tw).count());
  // the sketch entries are copied in by numeric() before step rng_step_
  rng_step_ = -1;
  rng_on_device_ = false;
  start_rng();
  // front descriptors
  std::vector<dev::FrontDev> fr(ns);
  {
    int co = 0;
    int64_t ro = 0;
    for (int j = 0; j < ns; ++j) {
      const Supernode& s = P.nodes[j];
      dev::FrontDev& f = fr[j];
      f.panel = s.compress ? (int64_t)((Wb() + s.off_work_panel) - F())
                           : s.off_panel;
      f.update = s.R() > 0 ? (int64_t)((Wb() + upd_base() + s.off_update) - F()) : 0;
      f.ncols = s.ncols;
      f.R = s.R();
      f.first = s.first;
      f.child_off = co;
      f.nchild = (int)s.children.size();
      co += f.nchild;
      f.relmap_off = ro;
      ro += (int64_t)s.rows.size();
    }
  }
  dfronts_.upload(fr, st_);
  double t_comp = 0;
  auto tl0 = std::chrono::steady_clock::now();
  for (size_t lv = 0; lv < P.levels.size(); ++lv) {
    const std::vector<int>& nodes = P.levels[lv];
    ctx_ = "level " + std::to_string(lv) + " uncompressed";
    Step st{OP_ASM, 0, (int64_t)tasks_.size()};
    for (int j : nodes) {
      const int w = P.nodes[j].ncols + P.nodes[j].R();
      for (int c0 = 0; c0 < w; c0 += 16) tasks_.push_back({j, c0});
    }
    st.count = (int)(tasks_.size() - st.first);
    if (st.count > 0) push(st);
    std::vector<CholMat> mats;
    std::vector<dev::GemmJob> syrk;
    for (int j : nodes) {
      const Supernode& s = P.nodes[j];
      if (s.compress) continue;
      const int c = s.ncols, R = s.R(), ld = c + R;
      mats.push_back({fp(s.off_panel), c, ld, ld, j, factor_tol()});
      if (R > 0) {
        dev::GemmJob g;
        g.A = F() + s.off_panel + c;
        g.B = F() + s.off_panel + c;
        g.C = Wb() + upd_base() + s.off_update;
        g.m = R;
        g.n = R;
        g.k = c;
        g.lda = g.ldb = ld;
        g.ldc = R;
        g.alpha = -1.0;
        g.beta = 1.0;
        syrk.push_back(g);
      }
    }
    panel_chol(mats);
    emit_gemm(OP_GEMM_NT, syrk, true);
    inv_cursor_ = inv_base();
    invert(mats);
    // Partitioned-inverse form of the uncompressed panels: Zp = L21 Lc^{-1}
    // replaces L21 once the Schur update has used it.  The forward solve then
    // needs only Lc^{-1} fv_c and fv_R -= Zp fv_c, both from fv_c; the
    // transposed solve x_c = Lc^{-T} y_c - Zp^T x_R.
    {
      ctx_ = "level " + std::to_string(lv) + " Zp";
      std::vector<dev::GemmJob> zg;
      std::vector<dev::CopyJob> zc;
      int64_t zo = zp_base();
      for (int j : nodes) {
        const Supernode& s = P.nodes[j];
        if (s.compress || s.R() == 0) continue;
        const int c = s.ncols, R = s.R(), ld = c + R;
        T* Z = Wb() + zo;
        zo += (int64_t)R * c;
        dev::GemmJob g;
        g.A = F() + s.off_panel + c;  // L21 (R x c)
        g.B = F() + s.off_panel;      // Lc^{-T} (upper), used transposed
        g.C = Z;
        g.m = R;
        g.n = c;
        g.k = c;
        g.lda = g.ldb = ld;
        g.ldc = R;
        g.alpha = 1.0;
        g.beta = 0.0;
        zg.push_back(g);
        zc.push_back({Z, F() + s.off_panel + c, R, c, R, ld});
      }
      if (zo - zp_base() > zp_elems_) throw std::runtime_error("Zp scratch too small");
      emit_gemm(OP_GEMM_NT, zg, false);
      emit_copy(zc);
    }
    std::vector<int> comp;
    for (int j : nodes)
      if (P.nodes[j].compress) comp.push_back(j);
    auto tc = std::chrono::steady_clock::now();
    emit_level_compressed(comp);
    t_comp += std::chrono::duration<double>(std::chrono::steady_clock::now() - tc).count();

  }
  if (P.opt.verbose)
    std::printf("program levels: %.3f s (compressed supernodes %.3f s), %zu steps, %zu GEMM "
                "tiles\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - tl0).count(),
                t_comp, prog_.size(), gtiles_.size());
  if (debug_check_) add_checks();
  dgjobs_.upload(gjobs_, st_);
  dgtiles_.upload(gtiles_, st_);
  dgpart_.alloc((size_t)std::max<int64_t>(gpart_bytes_, 8));
  dgcnt_.alloc((size_t)std::max(ngcnt_, 1) * sizeof(unsigned));
  RSC_HIP(hipMemsetAsync(dgcnt_.p, 0, dgcnt_.bytes, st_));
  dpjobs_.upload(pjobs_, st_);
  dtjobs_.upload(tjobs_, st_);
  dttiles_.upload(ttiles_, st_);
  dcjobs_.upload(cjobs_, st_);
  dctiles_.upload(ctiles_, st_);
  dsjobs_.upload(sjobs_, st_);
  dejobs_.upload(ejobs_, st_);
  dkjobs_.upload(kjobs_, st_);
  dvjobs_.upload(vjobs_, st_);
  detiles_.upload(etiles_, st_);
  dtasks_.upload(tasks_, st_);
  RSC_HIP(hipStreamSynchronize(st_));
  program_ready_ = true;
}

// One pass of the program: the launches of every step, in order.
template <typename T>
void Engine<T>::run_program(std::vector<hipEvent_t>* ev) {
  for (size_t si = 0; si < prog_.size(); ++si) {
    if (ev && si > 0) RSC_HIP(hipEventRecord((*ev)[si], st_));
    const Step& s = prog_[si];
    if ((int)si == rng_step_ && !rng_on_device_) upload_rng();
    switch (s.op) {
      case OP_ASM:
        dev::launch_assemble<T>(F(), dfronts_.as<dev::FrontDev>(),
                                dtasks_.as<dev::AsmTask>() + s.first, s.count,
                                dchildren_.as<int>(), drelmap_.as<int>(),
                                dlowp_.as<int>(), da_dst_.as<int64_t>(),
                                da_src_.as<int64_t>(), dAx_.as<double>(), st_);
        break;
      case OP_GEMM_NN:
      case OP_GEMM_TN:
      case OP_GEMM_NT:
      case OP_GEMM_TT: {
        bool ta = s.op == OP_GEMM_TN || s.op == OP_GEMM_TT;
        bool tb = s.op == OP_GEMM_NT || s.op == OP_GEMM_TT;
        if (s.f64)
          dev::launch_gemm<double>(ta, tb, dgjobs_.as<dev::GemmJob>(),
                                   dgtiles_.as<dev::GemmTile>() + s.first, s.count,
                                   dgpart_.p, dgcnt_.as<unsigned>(), st_);
        else if (s.bf16 && sizeof(T) == 4)
          dev::launch_gemm_bf16(ta, tb, dgjobs_.as<dev::GemmJob>(),
                                dgtiles_.as<dev::GemmTile>() + s.first, s.count, dgpart_.p,
                                dgcnt_.as<unsigned>(), st_);
        else
          dev::launch_gemm<T>(ta, tb, dgjobs_.as<dev::GemmJob>(),
                              dgtiles_.as<dev::GemmTile>() + s.first, s.count,
                              dgpart_.p, dgcnt_.as<unsigned>(), st_);
        break;
      }
      case OP_POTRF:
        if (s.f64)
          dev::launch_potrf<double>(dpjobs_.as<dev::PotrfJob>() + s.first, s.count,
                                    dstatus_.as<int>(), st_);
        else
          dev::launch_potrf<T>(dpjobs_.as<dev::PotrfJob>() + s.first, s.count,
                               dstatus_.as<int>(), st_);
        break;
      case OP_TRSM:
        if (s.f64)
          dev::launch_trsm<double>(dtjobs_.as<dev::TrsmJob>(),
                                   dttiles_.as<dev::RowTile>() + s.first, s.count, st_);
        else
          dev::launch_trsm<T>(dtjobs_.as<dev::TrsmJob>(),
                              dttiles_.as<dev::RowTile>() + s.first, s.count, st_);
        break;
      case OP_CONV:
        dev::launch_convjobs(dvjobs_.as<dev::ConvJob>() + s.first, s.count, st_);
        break;
      case OP_COPY:
        dev::launch_copy<T>(dcjobs_.as<dev::CopyJob>(),
                            dctiles_.as<dev::RowTile>() + s.first, s.count, st_);
        break;
      case OP_SHIFT:
        if (s.f64)
          dev::launch_shift<double>(dsjobs_.as<dev::ShiftJob>() + s.first, s.count, st_);
        else
          dev::launch_shift<T>(dsjobs_.as<dev::ShiftJob>() + s.first, s.count, st_);
        break;
      case OP_CHECK:
        if (s.f64)
          dev::launch_check<double>(dkjobs_.as<dev::CheckJob>() + s.first, s.count,
                                    dstatus_.as<int>(), st_);
        else
          dev::launch_check<T>(dkjobs_.as<dev::CheckJob>() + s.first, s.count,
                               dstatus_.as<int>(), st_);
        break;
      case OP_EYE:
        dev::launch_eye<T>(dejobs_.as<dev::EyeJob>(),
                           detiles_.as<dev::RowTile>() + s.first, s.count, st_);
        break;
    }
  }
}

template <typename T>
void Engine<T>::numeric(const double* Ax) {
  auto tv = std::chrono::steady_clock::now();
  auto vphase = [&](const char* what) {
    if (!plan_.opt.verbose) return;
    auto now = std::chrono::steady_clock::now();
    std::printf("numeric %s: %.3f s\n", what, std::chrono::duration<double>(now - tv).count());
    tv = now;
  };
  join_program_build();
  vphase("program build (joined)");
  if (!program_ready_) {
    drop_graph();
    build_program();
    vphase("program build");
  }
  // matrix values into a buffer that keeps its address (the graph reads it)
  const size_t abytes = std::max<size_t>((size_t)A_.nnz() * sizeof(double), 8);
  if (!dAx_.p || dAx_.bytes != abytes) dAx_.alloc(abytes);
  if (A_.nnz() > 0)
    RSC_HIP(hipMemcpyAsync(dAx_.p, Ax, (size_t)A_.nnz() * sizeof(double),
                           hipMemcpyHostToDevice, st_));
  int status0[4] = {0, INT_MAX, 0, INT_MAX};
  RSC_HIP(hipMemcpyAsync(dstatus_.p, status0, sizeof(status0),
                         hipMemcpyHostToDevice, st_));
  // optional timing trace: an event after every step (RSC_FACTOR_TRACE=path)
  const char* ftp = std::getenv("RSC_FACTOR_TRACE");
  std::vector<hipEvent_t> fev;
  if (ftp && ftp[0]) {
    fev.resize(prog_.size() + 1);
    for (auto& e : fev) RSC_HIP(hipEventCreate(&e));
    RSC_HIP(hipEventRecord(fev[0], st_));
  }
  // Later factorizations replay the program as a graph (RSC_FACTOR_GRAPH=0
  // keeps every launch eager).  The first one runs eagerly: it copies the
  // sketch entries in mid-program, while the host is still drawing them.
  const char* ge = std::getenv("RSC_FACTOR_GRAPH");
  const bool graph = !(ge && ge[0] == '0') && fev.empty() && !debug_check_ &&
                     !graph_failed_ && (rng_on_device_ || plan_.rng_count <= 0) &&
                     !prog_.empty();
  if (graph && !fgexec_) {
    hipError_t e = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      run_program(nullptr);
      e = hipStreamEndCapture(st_, &fgraph_);
      if (e == hipSuccess) e = hipGraphInstantiate(&fgexec_, fgraph_, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      drop_graph();
      graph_failed_ = true;
    }
  }
  if (graph && fgexec_) RSC_HIP(hipGraphLaunch(fgexec_, st_));
  else run_program(fev.empty() ? nullptr : &fev);
  RSC_HIP(hipGetLastError());
  if (!fev.empty()) RSC_HIP(hipEventRecord(fev[prog_.size()], st_));
  vphase("launches");
  int status[4];
  RSC_HIP(hipMemcpyAsync(status, dstatus_.p, sizeof(status), hipMemcpyDeviceToHost, st_));
  RSC_HIP(hipStreamSynchronize(st_));
  vphase("device program");
  if (!fev.empty()) write_factor_trace(ftp, fev);
  if (rng_on_device_ && rng_host_) {
    RSC_HIP(hipHostFree(rng_host_));
    rng_host_ = nullptr;
  }
  finish_solve_setup();
  vphase("solve setup");
  pivot_replacements_ = status[2];
  if (status[3] != INT_MAX)
    throw DeviceError("non-finite factor values first after program step " +
                      std::to_string(status[3]) + ": " +
                      (status[3] < (int)step_desc_.size() ? step_desc_[status[3]]
                                                          : std::string("?")));
  if (status[0] != 0)
    throw IndefiniteError("indefinite diagonal block in supernode " +
                          std::to_string(status[1]));
}

template <typename T>
int64_t Engine<T>::apply_scratch_doubles(int k) const {
  // each part rounded up to 32 doubles (256 bytes) so every part is aligned
  auto up = [](int64_t v) { return (v + 31) / 32 * 32; };
  const int64_t n = plan_.n;
  return up(fv_stride_ * k) + up(tv_stride_ * k) + up(part_stride_ * k) + 2 * up(n * k);
}

template <typename T>
void Engine<T>::apply_into(int k, const double* B, int ldb, double* X, int ldx,
                           hipStream_t s, double* scratch) {
  if (!solve_ready_) throw std::logic_error("apply before the factorization finished");
  auto up = [](int64_t v) { return (v + 31) / 32 * 32; };
  const int64_t n = plan_.n;
  double* fv = scratch;
  double* tv = fv + up(fv_stride_ * k);
  double* part = tv + up(tv_stride_ * k);
  double* w = part + up(part_stride_ * k);
  double* z = w + up(n * k);
  dev::VecBufs vb{w, fv, tv, part, n, fv_stride_, tv_stride_, part_stride_, nullptr, z, 2};
  run_apply(k, B, ldb, X, ldx, s, vb);
}

template <typename T>
void Engine<T>::apply(int k, const double* B, int ldb, double* X, int ldx,
                      hipStream_t s, const int* live, int multi_min_k) {
  finish_solve_setup();
  const int n = plan_.n;
  if (k > fv_k_) {
    dfv_.alloc((size_t)fv_stride_ * k * sizeof(double));
    dtv_.alloc((size_t)tv_stride_ * k * sizeof(double));
    dpart_.alloc((size_t)part_stride_ * k * sizeof(double));
    dw_.alloc((size_t)n * k * sizeof(double));
    dz_.alloc((size_t)n * k * sizeof(double));
    fv_k_ = k;
  }
  dev::VecBufs vb{dw_.as<double>(), dfv_.as<double>(), dtv_.as<double>(),
                  dpart_.as<double>(), (int64_t)n, fv_stride_, tv_stride_, part_stride_,
                  live, dz_.as<double>(), multi_min_k};
  run_apply(k, B, ldb, X, ldx, s, vb);
}

// The launches of one apply on given vector buffers (no allocation).
template <typename T>
void Engine<T>::run_apply(int k, const double* B, int ldb, double* X, int ldx,
                          hipStream_t s, const dev::VecBufs& vb) {
  const int n = plan_.n;
  const int multi_min_k = vb.multi_min_k;
  const int* live = vb.live;
  if (use_dataflow_) {
    dev::DfArgs a{};
    df_last_ = (k > 1 || multi_min_k <= 1) ? 1 : 0;
    DfList& D = df_[df_last_];
    a.tasks = D.tasks.as<dev::DfTask>();
    a.ntasks = D.ntasks;
    a.ngroups = D.ngroups;
    a.gv = D.gv.as<dev::GvTask>();
    a.gwait = D.gw.as<dev::DfWait>();
    a.cnt = D.cnt.as<unsigned>();
    a.status = dapst_.as<int>();
    a.timeout_ticks = df_timeout_;
    a.backoff_max = df_backoff_;
    a.nodes = dvnodes_.as<dev::VecNode>();
    a.children = dchildren_.as<int>();
    a.relmap = drelmap_.as<int>();
    a.ranges = dranges_.as<int>();
    a.rows = drows_.as<int>();
    a.perm = dperm_.as<int>();
    a.pinv = dpinv_.as<int>();
    a.B = B;
    a.ldb = ldb;
    a.X = X;
    a.ldx = ldx;
    a.trace = dtrace_.p ? dtrace_.as<unsigned long long>() : nullptr;
    a.sub = subtree_args(B, ldb, df_last_ == 1);
    if (df_queue_) {
      a.q_queue = D.q_queue.as<int>();
      a.q_ctr = D.q_ctr.as<unsigned>();
      a.q_pend = D.q_pend.as<int>();
      a.q_need = D.q_need.as<unsigned>();
      a.q_gptr = D.q_gptr.as<int>();
      a.q_gtasks = D.q_gtasks.as<int>();
      a.q_ginfo = D.q_ginfo.as<int4>();
      a.q_sptr = D.q_sptr.as<int>();
      a.q_succ = D.q_succ.as<int>();
      dev::launch_df_reset_queue(D.cnt.as<unsigned>(), D.q_pend.as<int>(), D.q_pend0.as<int>(),
                                 D.ngroups, D.q_queue.as<int>(), D.q_init.as<int>(), D.q_n0,
                                 D.ntasks, D.q_ctr.as<unsigned>(), s);
    } else {
      dev::launch_df_reset(D.cnt.as<unsigned>(), D.ngroups + 1, s);
    }
    dev::launch_apply_dataflow<T>(F(), a, vb, k, D.grid, s);
    return;
  }
  dev::launch_gather(B, ldb, vb.w, dperm_.as<int>(), n, k, live, s);
  const int nlev = (int)plan_.levels.size();
  auto run = [&](const std::vector<SolvePhase>& ph, int lv) {
    for (const SolvePhase& p : ph) {
      if (p.kind == 0)
        dev::launch_fwd_assemble(dvnodes_.as<dev::VecNode>(),
                                 dlevnodes_.as<int>() + lev_off_[lv],
                                 lev_off_[lv + 1] - lev_off_[lv], dchildren_.as<int>(),
                                 drelmap_.as<int>(), dranges_.as<int>(), vb, k,
                                 lev_front_[lv], s);
      else if (p.kind == 1)
        dev::launch_bwd_gather(dvnodes_.as<dev::VecNode>(),
                               dlevnodes_.as<int>() + lev_off_[lv],
                               lev_off_[lv + 1] - lev_off_[lv], drows_.as<int>(), vb, k,
                               lev_front_[lv], s);
      else if (p.kind == 4 || p.kind == 5)
        dev::launch_subtrees<T>(F(), subtree_args(B, ldb), p.count, p.kind == 4, vb, k, s);
      else
        dev::launch_gemv_tasks<T>(F(), dgv_.as<dev::GvTask>() + p.first, p.count, vb, k,
                                  p.kind == 3, s);
    }
  };
  for (int lv = 0; lv < nlev; ++lv) run(fwd_[lv], lv);
  for (int lv = nlev - 1; lv >= 0; --lv) run(bwd_[lv], lv);
  dev::launch_scatter(vb.w, X, ldx, dperm_.as<int>(), n, k, live, s);
}

template <typename T>
void Engine<T>::raise_apply_status(const int* st) {
  if (st[0] == 0 && st[1] == 0) return;
  // clear the words so the next apply starts clean, then report
  const int z[2] = {0, 0};
  RSC_HIP(hipMemcpy(dapst_.p, z, sizeof(z), hipMemcpyHostToDevice));
  throw DeviceError("preconditioner apply: a dependency wait timed out");
}

template <typename T>
void Engine<T>::check_apply(hipStream_t s) {
  if (!use_dataflow_) {
    RSC_HIP(hipStreamSynchronize(s));
    return;
  }
  int st[2] = {0, 0};
  RSC_HIP(hipMemcpyAsync(st, dapst_.p, sizeof(st), hipMemcpyDeviceToHost, s));
  RSC_HIP(hipStreamSynchronize(s));
  raise_apply_status(st);
  if (dtrace_.p) {
    // timing trace of the last dataflow apply: ntasks, tasks, 4 clocks each,
    // then the wait-list length, the wait list, the matrix-vector entries and
    // four cost-model values per task
    const DfList& DL = df_[df_last_];
    const std::vector<dev::DfTask>& df_host_ = DL.host;
    std::vector<unsigned long long> tr(4 * df_host_.size());
    RSC_HIP(hipMemcpy(tr.data(), dtrace_.p, tr.size() * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
    if (FILE* fp = std::fopen(df_trace_path_.c_str(), "wb")) {
      const int64_t nt = (int64_t)df_host_.size();
      std::fwrite(&nt, sizeof(nt), 1, fp);
      std::fwrite(df_host_.data(), sizeof(dev::DfTask), df_host_.size(), fp);
      std::fwrite(tr.data(), sizeof(unsigned long long), tr.size(), fp);
      const int64_t nw = (int64_t)DL.host_wait.size();
      std::fwrite(&nw, sizeof(nw), 1, fp);
      std::fwrite(DL.host_wait.data(), sizeof(dev::DfWait), DL.host_wait.size(), fp);
      std::fwrite(DL.host_gv.data(), sizeof(dev::GvTask), DL.host_gv.size(), fp);
      std::fwrite(DL.host_feat.data(), sizeof(float), DL.host_feat.size(), fp);
      std::fclose(fp);
    }
  }
}

}  // namespace

std::unique_ptr<FactorEngine> make_engine(const Plan& plan, const Csc& A,
                                          hipStream_t stream) {
  if (plan.opt.factor_precision == RSC_FP32)
    return std::unique_ptr<FactorEngine>(new Engine<float>(plan, A, stream));
  return std::unique_ptr<FactorEngine>(new Engine<double>(plan, A, stream));
}

}  // namespace rsc
