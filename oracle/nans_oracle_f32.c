/* Binary32 instantiation of the CPU oracle (test infrastructure only). */
#include "nans_oracle.h"
#define REAL float
#define SUF(x) x##_f32
#define ORA_F32 1
#include "nans_oracle_impl.h"
