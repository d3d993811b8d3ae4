"""CPU oracle for the separation likelihood.

TEST INFRASTRUCTURE ONLY: a C restatement of the reference's FP64 CPU path
(``mw_oracle.c``) and its ctypes wrapper (``oracle.py``).  Only tests/, the
smoke check and the CPU-baseline leg of bench.py may import it; the product
package never does.
"""
