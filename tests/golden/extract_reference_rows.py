"""Writes the small fixtures of this folder from the reference's recorded
outputs (run where the reference tree is present; the fixtures are
committed, the reference is not):

  tgv64_taylorgreen_out_rows1-2.txt    first two rows of
      taylorgreen_vortex/taylorgreen_run64/taylorgreen.out
  tgv128_taylorgreen_out_rows1-2.txt   first two rows of
      taylorgreen_vortex/taylorgreen_run128/taylorgreen.out
  hit128_k8_turbdata_out_rows1-2.txt   first two rows of
      homogenous_turbulence/Samtaney_D4/turbdata_N128_D4_k8.out
  tgv64_taylorgreen_out_steps.txt      the 64^3 rows after steps 1-10 and
      every tenth step to 1000, each preceded by its step number
  tgv128_taylorgreen_out_steps.txt     the 128^3 rows after steps 1-10 and
      every tenth step to 200, likewise

(row n of taylorgreen.out is written after step n; its first column is the
time at the start of that step)

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

STEPS = {
    "tgv64_taylorgreen_out_steps.txt": ("taylorgreen_vortex/taylorgreen_run64/taylorgreen.out",
                                        1000),
    "tgv128_taylorgreen_out_steps.txt": ("taylorgreen_vortex/taylorgreen_run128/taylorgreen.out",
                                         200),
}

for out, (src, last) in STEPS.items():
    keep = set(range(1, 11)) | set(range(10, last + 1, 10))
    with open(os.path.join(REF, src)) as f:
        lines = [line for _, line in zip(range(last), f)]
    with open(os.path.join(HERE, out), "w") as f:
        for n in sorted(keep):
            f.write(f"{n} {lines[n - 1]}")
