"""CPU oracle for the reconstruction hot path.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import it, and only as the checker (or the timed CPU
baseline).  The product never imports it.

  recon.py       NumPy restatement (FP64 accuracy mode, FP32 reference-order
                 mode) of every hot-path function, each citing the reference
                 file:line it follows.
  bp_oracle.c    C/OpenMP restatement of the FBP/FDK back-projector (second,
                 independent restatement; the CPU baseline).
  cbp.py         ctypes loader of libbp_oracle.so.
  build.py       gcc recipe for libbp_oracle.so.
"""
