// Batched persistent min-sum engine, FP16 message format.
#define BATCH_FMT 1
#define BATCH_LAUNCH batch_launch_f16
#define BATCH_CAPACITY batch_capacity_f16
#include "ldpc_batch_impl.h"
