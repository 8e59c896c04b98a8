"""oracle — TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's algorithm for the FP16 Winograd F(2x2,3x3) conv2d path
(Yhorm26/2024PRA_CONV). Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import it, and only as the checker: the product (the gfx950 HIP library) never calls
into it, and fails loudly when its own extension is missing.

Contents
  direct_conv.c / cdirect.py : FP64 direct convolution with exact products and the faithful
                               restatement of the reference checker conv2dcpu
                               (include/verfiy.h:5-43); ctypes bindings.
  winograd.py                : numpy model of the four Winograd stages in FP64 and in the
                               GPU's FP32 operation order, the packed U/V layouts, seeded
                               input generators and the error metrics.
  _ref/                      : the reference's own checker compiled from its source (built
                               where the reference is present; see Makefile).

Parity status of each function is listed in the module docstrings; every function is
pinned by at least one independent check in tests/test_oracle.py.
"""
