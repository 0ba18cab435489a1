"""The build's spill guard (3dcompactsolver_amd/build.py): the per-kernel
resource report of hipcc is parsed, and a kernel with scratch fails the
build.  Here on a report of the form hipcc prints, and on a real compile
of a kernel that spills and one that does not."""
import os
import subprocess

import pytest

import build

REPORT = """\
stage_z.hip:27:1: remark: Function Name: _ZN4cs3d7k_zpassIfLi4ELb0ELb1EE [-Rpass-analysis=kernel-resource-usage]
stage_z.hip:27:1: remark:     VGPRs: 128 [-Rpass-analysis=kernel-resource-usage]
stage_z.hip:27:1: remark:     ScratchSize [bytes/lane]: 80 [-Rpass-analysis=kernel-resource-usage]
stage_z.hip:27:1: remark: Function Name: _ZN4cs3d7k_zpassIfLi4ELb1ELb0EE [-Rpass-analysis=kernel-resource-usage]
stage_z.hip:27:1: remark:     VGPRs: 128 [-Rpass-analysis=kernel-resource-usage]
stage_z.hip:27:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]
"""


def test_report_parser():
    assert build.scratch_kernels(REPORT) == [("_ZN4cs3d7k_zpassIfLi4ELb0ELb1EE", 80)]
    assert build.scratch_kernels("") == []


SRC = """
#include <hip/hip_runtime.h>
template <int N>
__global__ __launch_bounds__(1024) void k(const float *a, float *b, int s)
{
    float r[N];
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = a[threadIdx.x + i * s];
    float acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) acc += r[i] * r[(i * 7 + j) %% N];
#pragma unroll
    for (int i = 0; i < N; i++) b[threadIdx.x + i * s] = r[i] + acc;
}
template __global__ void k<%d>(const float *, float *, int);
"""


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="hipcc not installed")
@pytest.mark.parametrize("n,spills", [(16, False), (200, True)])
def test_guard_on_real_compile(tmp_path, n, spills):
    src = tmp_path / "k.hip"
    src.write_text(SRC % n)
    r = subprocess.run([build.HIPCC] + build.COMMON +
                       ["-c", str(src), "-o", str(tmp_path / "k.o"), build.RESOURCE_REPORT],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert bool(build.scratch_kernels(r.stderr)) == spills
