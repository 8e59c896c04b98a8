// oracle/ref_driver.cpp — TEST INFRASTRUCTURE ONLY.
// Exposes the reference's own CPU checker conv2dcpu (include/verfiy.h:5-43, compiled from
// the reference source where it lies) with C linkage, so tests can pin the oracle against
// it and bench.py can time it as the CPU baseline.
#include "verfiy.h"

extern "C" void ref_conv2dcpu(_Float16* pin, _Float16* pwei, _Float16* pout, int n, int c,
                              int h, int w, int k, int r, int s, int u, int v, int p, int q) {
    conv2dcpu(pin, pwei, pout, n, c, h, w, k, r, s, u, v, p, q);
}
