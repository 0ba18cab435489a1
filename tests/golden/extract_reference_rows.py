"""Writes the small fixtures of this folder from the reference's recorded
outputs (run where the reference tree is present; the fixtures are
committed, the reference is not):

  tgv64_taylorgreen_out_rows1-2.txt    first two rows of
      taylorgreen_vortex/taylorgreen_run64/taylorgreen.out
  tgv128_taylorgreen_out_rows1-2.txt   first two rows of
      taylorgreen_vortex/taylorgreen_run128/taylorgreen.out
  hit128_k8_turbdata_out_rows1-2.txt   first two rows of
      homogenous_turbulence/Samtaney_D4/turbdata_N128_D4_k8.out

usage: python extract_reference_rows.py [reference checkout]
"""
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FIXTURES = {
    "tgv64_taylorgreen_out_rows1-2.txt": "taylorgreen_vortex/taylorgreen_run64/taylorgreen.out",
    "tgv128_taylorgreen_out_rows1-2.txt": "taylorgreen_vortex/taylorgreen_run128/taylorgreen.out",
    "hit128_k8_turbdata_out_rows1-2.txt":
        "homogenous_turbulence/Samtaney_D4/turbdata_N128_D4_k8.out",
}

for out, src in FIXTURES.items():
    with open(os.path.join(REF, src)) as f:
        rows = [next(f) for _ in range(2)]
    with open(os.path.join(HERE, out), "w") as f:
        f.writelines(rows)
