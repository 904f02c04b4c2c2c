"""ORACLE -- test infrastructure only.

A CPU restatement of the mbirjax 0.6.7 projector + VCD hot path.  It exists to
check the MI355X kernels (mbirjax_amd) and is imported only by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg; the product never
imports, links or executes anything under oracle/.

Modules:
  projectors  FP64 (and FP32-port) C restatement of the sparse forward/back
              projectors (mbx_oracle.c) plus the geometry helpers
  qggmrf      NumPy FP64 restatement of mbirjax/qggmrf.py
  partitions  NumPy restatement of the pixel partitions (mbirjax/vcd_utils.py)
  filters     FBP/FDK filtering (mbirjax/tomography_utils.py, cone_beam.py, parallel_beam.py)
  vcd         FP64 VCD reconstruction (mbirjax/tomography_model.py:756-1658)

Parity status of each function is recorded in the module headers.
"""
