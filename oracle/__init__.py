"""CPU oracle for the reffine scan path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import anything under oracle/, and only as the checker (or the timed CPU
baseline); the product path (reffine_amd/) never calls it.

  oracle.brute      brute-force restatement of each query (exact integers,
                    correctly rounded FP64 sums via math.fsum)
  oracle.detorder   the deterministic reduction order of the GPU kernels,
                    restated in numpy for bit-exact comparison
  oracle.cjit       CPU JIT restatement of reffine's LoopGen -> CEmitter ->
                    clang -O3 path (the reported CPU baseline)
"""
