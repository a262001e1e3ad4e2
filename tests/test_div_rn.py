"""div_rn (so3g_amd/csrc/ptg.h), the division by a correctly rounded
reciprocal that the pointing kernels use in place of a / b: built for the
host with hipcc from the same header and compared bit for bit with IEEE
division (tests/native/div_rn_check.cpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


@pytest.mark.skipif(not os.path.exists(HIPCC), reason='hipcc not installed')
def test_div_rn_matches_division():
    src = os.path.join(HERE, 'native', 'div_rn_check.cpp')
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, 'div_rn_check')
        subprocess.run([HIPCC, '-O2', '-std=c++17', '-x', 'hip', '--offload-arch=gfx950', src,
                        '-o', exe], check=True, capture_output=True)
        out = subprocess.run([exe, '1000000'], capture_output=True, text=True, timeout=600)
    checks, bad = (int(v) for v in out.stdout.split()[-2:])
    assert checks > 7e6
    assert bad == 0, out.stdout
