// batch.h -- batched launch helpers shared by batch.hip and serve.hip.
#pragma once
#include "common.h"

namespace sable {
int batch_spmv_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int nrhs, const sable_batch_spmv_item_t* items,
                      int nitems, int64_t total_blocks, unsigned long long* counter, int wgs);
int batch_spts_launch(hipStream_t st, int num_cu, sable_dtype_t dtype, int lanes, const sable_batch_spts_item_t* items,
                      int nitems, int64_t total_rows, unsigned* ticket, int* err, long long spin);
int batch_refine_launch(hipStream_t st, int num_cu, int lanes, const sable_batch_refine_item_t* items,
                        sable_batch_spts_item_t* spts_view, int nitems, int64_t total_rows, int iters,
                        unsigned* ticket, int* err, long long spin);
}  // namespace sable
