"""The oracle: a CPU (numpy) restatement of the reference's CRV response chain.

Test infrastructure only. Tests, the smoke check and the CPU-baseline leg of the benchmark use it as the checker;
the product path (crvhip, libcrvhip.so) never imports or calls it.
"""
