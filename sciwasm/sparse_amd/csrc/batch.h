// batch.h -- batched launch helpers shared by batch.hip and serve.hip.
#pragma once
#include "common.h"

#include <vector>

namespace sable {
int batch_spmv_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int nrhs, const sable_batch_spmv_item_t* items,
                      int nitems, int64_t total_blocks, unsigned long long* counter, int wgs);
int batch_spts_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                      int nitems, int64_t total_rows, unsigned* ticket, int* err, long long spin);
int batch_refine_launch(hipStream_t st, int num_cu, int lanes, const sable_batch_refine_item_t* items,
                        sable_batch_spts_item_t* spts_view, int nitems, int64_t total_rows, int iters,
                        unsigned* ticket, int* err, long long spin);

// Work of the server's batched solves, cut on the host into chunks of at
// most one workgroup's rows (256 / lanes), so that no kernel searches for
// the item a row belongs to.  rows != nullptr: the chunk's rows are
// rows[0..count) of the item (level-ordered row lists); otherwise rows
// first .. first + count - 1.
struct RowChunk {
    const int32_t* rows;
    int32_t item;
    int32_t first;
    int32_t count;
    int32_t pad;
};
// Level l of a batch is chunks [chunk_lo[l], chunk_lo[l + 1]).
struct LevelPlan {
    std::vector<int32_t> chunk_lo;
    int levels() const { return chunk_lo.empty() ? 0 : (int)chunk_lo.size() - 1; }
};
// Every item solved level by level, one launch per level (all items' rows
// of level l in launch l); x is written, never read before it is.
int batch_level_launch(hipStream_t st, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                       const RowChunk* d_level_chunks, const LevelPlan& plan);
// FP32 level-set solves with `iters` FP64 residual + update steps; the
// residual and update run over d_row_chunks (natural row order).  One step
// (iters = 1) is fused into one sweep of the level chunks: every row forms
// its FP32 x0, FP64 residual, FP32 correction and x in the same launch.
int batch_refine_level_launch(hipStream_t st, int lanes, const sable_batch_refine_item_t* items,
                              sable_batch_spts_item_t* spts_view, int nitems, int iters,
                              const RowChunk* d_level_chunks, const LevelPlan& plan, const RowChunk* d_row_chunks,
                              int nrow_chunks);
}  // namespace sable
