"""Host-side checks of bench.py's contract that need no GPU: the one-GPU
refusal for --gpus above 1 and the measured-traffic lookup from the
committed rocprofv3 PMC summaries under profiles/."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_multi_gpu_refused_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'),
                        '--gpus', '2', '--steps', '1', '--warmup', '0'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert 'one GPU' in line['error']


def test_measured_traffic_per_launch():
    # every kernel of the direct pipeline has a traffic line, positive and
    # parsed as bytes per launch; template arguments after the name are
    # ignored, unknown kernels give None
    for k in ('batch_lines_thin', 'sep_xpass<2, 32>', 'sep_ypass<2, 32>',
              'sep_zpass<1, 12>', 'fx_rows_conv<3, 256, 300>',
              'fft_convolution', 'model_deposit', 'mcdisk_deposit'):
        t = bench.measured_traffic(k)
        assert t is not None and t > 0, k
    assert bench.measured_traffic('no_such_kernel') is None


def test_traffic_lines_well_formed():
    d = os.path.join(ROOT, 'profiles')
    seen = 0
    names = set()
    for fn in sorted(os.listdir(d)):
        if not fn.endswith('_traffic.txt'):
            continue
        with open(os.path.join(d, fn)) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch '):
                    name, val = line[len('traffic_bytes_per_launch '):].split(
                        ':', 1)
                    assert ' ' not in name.strip()
                    float(val)
                    # one summary per kernel: the lookup takes the first
                    # file's line, so a second one would be silently stale
                    assert name.strip() not in names, (fn, name)
                    names.add(name.strip())
                    seen += 1
    assert seen > 0
