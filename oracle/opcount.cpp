/*
 * opcount.cpp -- TEST INFRASTRUCTURE ONLY.  The op-counting build of the CPU oracle: the oracle's
 * own source compiled with every `float` and `double` replaced by the counting types of
 * opcount.h, exported as oracle/liboracle_opcount.so with the oracle's C entry points plus the
 * counter read-out.  Run single-threaded (the counters are per thread).  Used only to freeze the
 * floating-point work per Newton iteration (tests/golden/make_flop_counts.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <stdio.h>
#include <vector>
#include <algorithm>
#include <thread>
#include <atomic>
#include <mutex>
#include <functional>

#include "opcount.h"

namespace opc { thread_local uint64_t counts[2][NCOUNT]; }

#define float CFloat
#define double CDouble
#include "sms_oracle.cpp"
#undef float
#undef double

extern "C" {
/* out: [float add, mul, div, fma, sqrt, misc, double add, mul, div, fma, sqrt, misc] of this thread */
void opcount_read(uint64_t *out) {
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < opc::NCOUNT; ++i) out[k * opc::NCOUNT + i] = opc::counts[k][i];
}
void opcount_reset(void) { memset(opc::counts, 0, sizeof(opc::counts)); }
}
