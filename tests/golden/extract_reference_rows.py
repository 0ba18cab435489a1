"""Writes the small fixtures of this folder from the reference's recorded
outputs (run where the reference tree is present; the fixtures are
committed, the reference is not):

  hit128_k8_turbdata_out_rows1-2.txt   first two rows of
      homogenous_turbulence/Samtaney_D4/turbdata_N128_D4_k8.out
"""
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(REF, "homogenous_turbulence/Samtaney_D4/turbdata_N128_D4_k8.out")) as f:
    rows = [next(f) for _ in range(2)]
with open(os.path.join(HERE, "hit128_k8_turbdata_out_rows1-2.txt"), "w") as f:
    f.writelines(rows)
