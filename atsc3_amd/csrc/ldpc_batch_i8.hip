// Batched persistent min-sum engine, INT8 message format.
#define BATCH_FMT 2
#define BATCH_LAUNCH batch_launch_i8
#define BATCH_CAPACITY batch_capacity_i8
#include "ldpc_batch_impl.h"
