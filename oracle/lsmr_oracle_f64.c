/* lsmr_oracle_f64.c -- test infrastructure (the oracle): the LSMR
 * restatement of lsmr_oracle.c with every vector and scalar in double, a
 * yardstick for the float iterates of the reference and of the GPU solver;
 * not a restatement of any reference build. */
#define LSMR_F64
#include "lsmr_oracle.c"
