/* Binary64 instantiation of the CPU oracle (test infrastructure only). */
#include "nans_oracle.h"
#define REAL double
#define SUF(x) x##_f64
#define ORA_F32 0
#include "nans_oracle_impl.h"
