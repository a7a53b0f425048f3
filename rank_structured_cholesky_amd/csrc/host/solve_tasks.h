// Task lists of the preconditioner apply M^{-1} = P^T L^{-T} L^{-1} P, built
// on the host from the symbolic plan (FactorManager::forwardSolve /
// backwardSolve, FactorManager.cpp:6245-6341; Supernode::forwardSolve /
// backwardSolve, Supernode.cpp:4571-5250).  The structs in namespace dev are
// read by the device kernels as they are laid out here.
#pragma once

#include <cstdint>
#include <vector>

#include "rsc_host.h"

namespace rsc {
namespace dev {

// One matrix-vector task of a solve phase.  ROW: y[i] (op)= sum_j M[i,j] x[j]
// for i < rows (one thread per row).  COL: y[j] (op)= sum_i M[i,j] x[i] for
// j < cols (waves reduce over rows).  M is the factor slice at `mat`
// (leading dimension ld); x / y live in w (stride n), the front vectors fv
// (stride fvs) or the rank-sized buffers tv (stride tvs): sel 0 / 1 / 2.
// A COL task over many rows is split into row chunks that write partial sums
// (part >= 0: slot of 32 partials in the partial buffer); a reduce task then
// adds a group's partials in fixed order and applies op to y.  In the
// dataflow list a split COL task that sets a rank-sized vector has no reduce
// task: the ROW tasks that read that vector sum the partials themselves
// (xpart >= 0), in the reduce task's order.
struct GvTask {
  int64_t mat;
  int ld, rows, cols;
  int kind;  // 0 ROW, 1 COL, 2 REDUCE (part = first slot, rows = slot count,
             // ld = slot stride, 0 read as 1)
             // part >= 0 on ROW / COL: write partial sums to slot part (COL)
             // or slots part (rows 0..31) and part + 1 (rows 32..63) (ROW)
  int op;    // 0 set, 1 subtract
  int xsel, ysel;
  int part;
  int xpart;  // ROW, >= 0: x[j] = sum over s < xnsub of the partial slot
  int xnsub;  //   xpart + (j / 32) * xnsub + s, entry j % 32 (x itself unused)
  int64_t xoff, yoff;
};

// Per-supernode vector gathers of a solve level.
struct VecNode {
  int first, ncols, R, nchild;
  int child_off;
  int range_off;     // first (begin, end) pair of this node's assembly chunks
                     // in asm_ranges: chunk q, child i -> pair range_off + q*nchild + i
  int64_t fv_off;    // front vector offset (per RHS column)
  int64_t rows_off;  // row list offset
  int64_t relmap_off;
};


// The whole application M^{-1} B as ONE persistent launch.  Tasks are listed
// in a topological order and handed out through a ticket counter; every task
// belongs to a group (the tasks of one stage of one supernode: its
// assembly, one matrix-vector phase, one reduction) and waits until each
// group on its wait list has completed all of its tasks.  A workgroup only
// ever waits on groups whose tasks hold smaller tickets, which are already
// running or done, so the launch drains whatever the occupancy.
// Task t of the list is matrix-vector task gv[t] when its type is 2 (the
// GvTask array runs parallel to the task list), so a workgroup fetches both
// descriptors of its next task at once.
struct DfTask {
  int type;    // 0 forward assembly chunk, 1 backward gather chunk, 2 matrix-vector
               // task, 3 scatter chunk, 4 forward subtree, 5 backward subtree
  int a, b;    // type 0/1: node, chunk; type 2: own index; type 3: chunk; 4/5: subtree
  int group;   // group signalled on completion
  int group2;  // second group signalled (-1: none)
  int wbeg, wend;  // the group's waits: entries [wbeg, wend) of the wait list
  int wg0;         // first wait, inline: group (-1: none)
  unsigned wn0;    // and the task count that completes it
  int tag;     // (supernode + 1) << 8 | stage, for traces (supernode -1: scatter;
               // stage 0 assembly, 1 gather, 2 scatter, 3 / 4 subtree forward /
               // backward, 10+k chain A, 20+k chain B, 30+k chain C)
  unsigned need1, need2;  // task counts completing group and group2 (0: none),
                          // read with the descriptor for the completion step
  int succ1[2], succ2[2];  // groups waiting on group / group2: ready-queue
                           // successor list ranges [begin, end)
  int sfirst[2];           // the first entry of each of those lists (-1: empty)
};

// Ready-queue completion: a task of a group of at most kQDirectMax tasks
// counts down the pending counts of the groups waiting on it directly; in
// larger groups only the task completing the group does (one count-down per
// group instead of one per task).
constexpr unsigned kQDirectMax = 4;

// One entry of the wait list: a group and the task count that completes it.
struct DfWait {
  int group;
  unsigned need;
};

}  // namespace dev

constexpr int kSolveChunk = 256;     // front-vector entries per assembly/gather task
constexpr int kRowSplitCols = 1024;  // dataflow ROW tasks: columns per partial task
constexpr int kColRowsMax = 2048;    // dataflow COL tasks: most rows (x staged in LDS)
constexpr int kScatterChunk = 2048;  // vector entries per scatter task
// Subtree tasks: one workgroup runs every supernode of a small subtree in
// turn; a node's ncols and off-diagonal rows must fit its LDS vectors.
constexpr int kSubtreeMaxCols = 256;
constexpr int kSubtreeMaxRows = 512;
constexpr int kSubtreeRhs = 4;  // right-hand sides per pass over a subtree's panels
// Subtree tasks of the multi-right-hand-side apply: a supernode's whole
// front vector (ncols + R entries) and y (ncols entries) for sixteen
// right-hand sides live in LDS, so these limits are per supernode.
constexpr int kMsMaxCols = 32;
constexpr int kMsMaxFront = 224;
// Larger supernodes of such subtrees keep fv_c and y in LDS and fv_R in
// device memory (any R): at most kMlMaxCols columns.
constexpr int kMlMaxCols = 128;
// A multi-right-hand-side subtree task keeps its processed supernodes on a
// stack of this many entries (a supernode's children are its top entries).
constexpr int kMsStack = 64;
// A supernode of a multi-right-hand-side subtree task: uncompressed, and
// either small (front in LDS) or at most kMlMaxCols columns.
inline bool multi_subtree_node(int ncols, int R) {
  return (ncols <= kMsMaxCols && ncols + R <= kMsMaxFront) || ncols <= kMlMaxCols;
}

// Subtrees of uncompressed supernodes handled by one task each.
struct Subtrees {
  std::vector<int> of;          // per supernode: subtree task id, -1 if none
  std::vector<int> ptr, nodes;  // nodes of task s: each tree in depth-first
                                // postorder (a node follows its subtree)
  std::vector<int> root;        // first root supernode of task s
  std::vector<int> rptr, roots;  // every root of task s: roots[rptr[s] .. rptr[s + 1])
  std::vector<int> cols;        // per task: first column, column count
                                // (-1 when the columns are not contiguous)
};
// Maximal subtrees of uncompressed supernodes with at most max_cols columns
// and max_rows off-diagonal rows (and ncols + R <= max_front) each, whose
// panels hold at most max_elems elements in total; one task per subtree.
// multi: the supernode rule of the multi-right-hand-side list
// (multi_subtree_node) instead of kSubtreeMaxCols / kSubtreeMaxRows.
// max_tasks > 0: while there are more subtrees than that, the two smallest
// tasks are merged into one (a task then runs several trees in turn), so
// every subtree task of the list starts at once on max_tasks workgroups.
void build_subtrees(const Plan& P, int64_t max_elems, Subtrees& S, bool multi = false,
                    int max_tasks = 0);

struct SolvePhase {
  int kind;  // 0 forward assemble, 1 backward gather, 2 small tasks, 3 large tasks,
             // 4 forward subtree tasks, 5 backward subtree tasks
  int first, count;
};

// Vector layout of the solves: per supernode its front vector (ncols + R
// entries per right-hand side) and, for compressed nodes, rank-sized vectors
// for the off-diagonal pair and every HODLR node.
struct SolveLayout {
  // per front-vector chunk of kSolveChunk entries and per child: the child
  // rows [begin, end) whose parent positions fall in the chunk (relative
  // maps are increasing, so they are contiguous)
  std::vector<int> asm_ranges;
  std::vector<dev::VecNode> vn;
  std::vector<int> children, relmap, rows;
  std::vector<int64_t> fv_off, tv_node;
  std::vector<std::vector<int64_t>> tv_h;
  std::vector<int64_t> panel;  // factor offset of each uncompressed panel (-1 otherwise)
  int64_t fv_stride = 1, tv_stride = 1;
  Subtrees sub;        // subtree tasks of the single-right-hand-side list
  Subtrees sub_multi;  // and of the multi-right-hand-side list
};

struct LevelTasks {
  std::vector<dev::GvTask> gv;
  std::vector<std::vector<SolvePhase>> fwd, bwd;  // per level
  std::vector<std::vector<int>> level_nodes;      // per level, nodes outside subtrees
  int64_t slots = 0;                              // partial-sum slots (32 each)
};

struct DataflowTasks {
  std::vector<dev::DfTask> tasks;  // topological order
  std::vector<dev::GvTask> gv;     // one per task (used for type 2)
  std::vector<unsigned> need;      // tasks per group
  std::vector<int> gwoff, gw;      // groups waited on, per group
  int64_t slots = 0;
  double est_span_us = 0;          // makespan of the list schedule (0: level order)
  std::vector<float> trace_feat;   // per task: estimated us, and for subtree
                                   // tasks supernodes, front entries, kB
  // Ready-queue form of the same graph: a group's tasks are pushed to the
  // queue once every group it waits on has completed.
  std::vector<int> q_gptr, q_gtasks;  // tasks of group g: q_gtasks[q_gptr[g] ..)
  std::vector<int> q_ginfo;           // per group, one 16-byte record: q_gptr[g],
                                      // its task count, its first task (-1: none), 0
  std::vector<int> q_sptr, q_succ;    // groups waiting on group g: q_succ[q_sptr[g] ..)
  std::vector<int> q_pend0;           // per group: count-downs it waits for (each group
                                      // waited on: its task count up to kQDirectMax, else 1)
  std::vector<int> q_init;            // tasks ready at the start, in list order
};

// Vector layout; subtree_elems = 0 disables the single-right-hand-side
// subtree tasks, multi_subtree_elems = 0 the multi-right-hand-side ones.
// multi_max_tasks: subtree tasks of the multi-right-hand-side list at most
// (0: no merging).
void build_solve_layout(const Plan& P, SolveLayout& L, int64_t subtree_elems,
                        int64_t multi_subtree_elems = 0, int multi_max_tasks = 0);
void build_level_tasks(const Plan& P, const SolveLayout& L, LevelTasks& out);
// Ticket order of a dataflow list.  workers = 0 keeps the level order;
// otherwise the list is ordered by a list schedule on that many workgroups
// (resident workgroups of the launch), with task costs estimated for nrhs
// right-hand sides and a factor of elem_bytes per element.
// Consecutive kSolveChunk chunks per assembly / gather task of the
// single-right-hand-side list (its tasks are short: fewer, longer ones pay
// the per-task overhead less often); the multi-right-hand-side list has one.
constexpr int kSingleChunkGroup = 4;
struct TicketModel {
  int workers = 0;
  int nrhs = 1;
  int chunk_group = 1;  // chunks per assembly / gather task (b = group index)
  int elem_bytes = 8;
  int col_rows = 0;  // rows per COL task before partial sums (0: 512)
  int row_cols = 0;  // columns per ROW task before partial sums (0: kRowSplitCols)
  int64_t wide_elems = 0;  // products this large are split finer (0: 2^20 elements,
                           // and only for nrhs > 1)
};
// subtrees: the subtree tasks of the list (L.sub for the single-right-hand-
// side list, L.sub_multi for the multi-right-hand-side list), or null for
// every supernode as ordinary tasks
void build_dataflow_tasks(const Plan& P, const SolveLayout& L, DataflowTasks& out,
                          const Subtrees* subtrees, const TicketModel& tm = TicketModel());

}  // namespace rsc
