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

// Batched level-set solve.  One level of one item: `count` rows of the item,
// listed (natural row indices) at rows[0..count); `prefix` is the number of
// rows of the same level in earlier items.
struct LevelSeg {
    const int32_t* rows;
    int64_t prefix;
    int32_t item;
    int32_t count;
};
// Level l of the batch owns segments [seg_lo[l], seg_lo[l+1]) and rows[l] rows.
struct LevelPlan {
    std::vector<int32_t> seg_lo;
    std::vector<int64_t> rows;
    int levels() const { return (int)rows.size(); }
};
// Every item solved level by level, one launch per level (all items' rows
// of level l in launch l); x is written, never read before it is.
int batch_level_launch(hipStream_t st, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                       const LevelSeg* d_segs, const LevelPlan& plan);
// FP32 level-set solves with `iters` FP64 residual + update steps.
int batch_refine_level_launch(hipStream_t st, int lanes, const sable_batch_refine_item_t* items,
                              sable_batch_spts_item_t* spts_view, int nitems, int64_t total_rows, int iters,
                              const LevelSeg* d_segs, const LevelPlan& plan);
}  // namespace sable
