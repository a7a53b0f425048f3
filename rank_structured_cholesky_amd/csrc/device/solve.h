// Device kernels of the preconditioner apply and of PCG.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor.h"
#include "../host/solve_tasks.h"

namespace rsc {
namespace dev {

// w[i + j*n] = b[perm[i] + j*ldb]
void launch_gather(const double* b, int ldb, double* w, const int* perm, int n,
                   int k, const int* live, hipStream_t s);
// x[perm[i] + j*ldx] = w[i + j*n]
void launch_scatter(const double* w, double* x, int ldx, const int* perm, int n,
                    int k, const int* live, hipStream_t s);

struct VecBufs {
  double* w;        // permuted vector (y, then x), stride n per column
  double* fv;       // front vectors
  double* tv;       // rank-sized vectors
  double* part;     // partial sums of split column tasks
  int64_t n, fvs, tvs, ps;
  const int* live;  // when non-null and *live == 0 every kernel of the apply returns
  double* z;        // z = L^{-T} y per supernode, stride n per column
  // the dataflow apply runs the matrix-core variant (sixteen right-hand
  // sides per matrix pass) for k > 1, and for k = 1 when multi_min_k <= 1
  int multi_min_k = 2;
};

// forward: fv_c = w[cols], fv_R = 0, plus children's fv_R via relmap
void launch_fwd_assemble(const VecNode* nodes, const int* level_nodes, int count,
                         const int* children, const int* relmap, const int* ranges,
                         VecBufs b, int k, int max_front, hipStream_t s);
// backward: fv_R[a] = w[rows[a]], w[cols] = z[cols]
void launch_bwd_gather(const VecNode* nodes, const int* level_nodes, int count,
                       const int* rows, VecBufs b, int k, int max_front, hipStream_t s);
template <typename T>
void launch_gemv_tasks(const T* factor, const GvTask* tasks, int count, VecBufs b,
                       int k, bool big, hipStream_t s);

// ---- dataflow apply (tasks: DfTask) --------------------------------------
// Subtree tasks (one workgroup per subtree of small uncompressed supernodes).
struct SubtreeArgs {
  const int* ptr;        // nodes of subtree s: nodes[ptr[s] .. ptr[s+1])
  const int* nodes;
  const int* cols;       // subtree s: first column cols[2s], count cols[2s+1] (-1: scattered)
  const int64_t* panel;  // factor offset of every uncompressed panel
  const VecNode* vn;
  const int* children;
  const int* relmap;
  const int* rows;
  const int* perm;
  const double* B;       // right-hand sides (forward input)
  int64_t ldb;
};
template <typename T>
void launch_subtrees(const T* factor, const SubtreeArgs& a, int count, bool forward,
                     VecBufs b, int k, hipStream_t s);

struct DfArgs {
  const DfTask* tasks;
  int ntasks, ngroups;
  const GvTask* gv;
  const DfWait* gwait;   // wait list (DfTask::wbeg / wend index it)
  unsigned* cnt;         // ngroups completion counters, then the ticket counter
  int* status;           // [0] = 1 after a wait timed out, 2 on a task the kernel
                         // variant does not run; [1] = abort request
  unsigned long long timeout_ticks;  // wall-clock ticks before a wait gives up
  int backoff_max;       // longest pause between polls of a wait, in s_sleep(4) units
  const VecNode* nodes;
  const int* children;
  const int* relmap;
  const int* ranges;  // child row ranges per assembly chunk (VecNode::range_off)
  const int* rows;
  const int* perm;
  const int* pinv;  // inverse permutation: x[j] = w[pinv[j]] (final scatter)
  const double* B;
  int64_t ldb;
  double* X;
  int64_t ldx;
  unsigned long long* trace;  // optional: per ticket grab / ready / done clock, workgroup
  SubtreeArgs sub;
  // Ready-queue dispatch (when q_queue is set) instead of tickets: a
  // workgroup pops the next queue slot and runs the task written there; the
  // task completing a group's last wait pushes that group's tasks.
  int* q_queue;          // ntasks slots, -1 until written
  unsigned* q_ctr;       // [0] next slot to pop, [1] next slot to push
  int* q_pend;           // per group: waits not yet complete
  const unsigned* q_need;  // per group: tasks signalling it
  const int* q_gptr;     // tasks of group g: q_gtasks[q_gptr[g] .. q_gptr[g + 1])
  const int* q_gtasks;
  const int4* q_ginfo;   // per group: (q_gptr[g], its task count, its first task, 0)
  const int* q_sptr;     // groups waiting on g: q_succ[q_sptr[g] .. q_sptr[g + 1])
  const int* q_succ;
};

// Zero the counters (ngroups + 1 entries) before a launch.
void launch_df_reset(unsigned* cnt, int count, hipStream_t s);
// Ready-queue state before a launch: counters zero, pend = pend0, the first
// n0 slots the initial tasks, the others empty (-1), pop 0 / push n0.
void launch_df_reset_queue(unsigned* cnt, int* pend, const int* pend0, int ngroups,
                           int* queue, const int* init, int n0, int ntasks, unsigned* ctr,
                           hipStream_t s);
template <typename T>
void launch_apply_dataflow(const T* factor, const DfArgs& a, VecBufs b, int k, int grid,
                           hipStream_t s);
// Workgroups of the dataflow kernel resident per compute unit.
template <typename T>
int dataflow_blocks_per_cu(bool multi);

// CSR (= CSC of the symmetric matrix) y = A x for k columns.
void launch_spmv(int n, const int* p, const int* idx, const double* val,
                 const double* x, int ldx, double* y, int ldy, int k,
                 const int* active, hipStream_t s);

}  // namespace dev
}  // namespace rsc
