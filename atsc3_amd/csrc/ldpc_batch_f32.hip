// Batched persistent min-sum engine, FP32 message format.
#define BATCH_FMT 0
#define BATCH_LAUNCH batch_launch_f32
#define BATCH_CAPACITY batch_capacity_f32
#include "ldpc_batch_impl.h"
