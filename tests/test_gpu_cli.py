"""The command-line driver (reference grammar) on the GPU: it compresses a raw
file, reports CR / PSNR / the random-access query, and its saved stream and
decoded output agree with the library called directly."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "sc25-stz2_amd", "lib", "stz_cli")


def _val(out, key):
    m = re.search(re.escape(key) + r"\s*([-+0-9.eE]+)", out)
    assert m, (key, out)
    return float(m.group(1))


@pytest.mark.parametrize("dt,flag,mode", [(np.float32, "-f", "--rel"), (np.float64, "-d", None)])
def test_cli_round_trip(stz, ctx, tmp_path, dt, flag, mode):
    from stz_amd import fields
    f = fields.sinusoids(shape=(40, 48, 56), seed=21, dtype=dt)
    raw = tmp_path / "in.raw"
    f.tofile(raw)
    out = tmp_path / "out.raw"
    stream = tmp_path / "s.stz"
    eb = 1e-3 if mode else 1e-3 * float(f.max() - f.min()) / 8.2
    cmd = [CLI, repr(eb), str(raw), "56", "48", "40", flag, "3", "4", "5", "40", "30", "20",
           "--out", str(out), "--save-stream", str(stream)] + ([mode] if mode else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cr = _val(r.stdout, "CR:")
    psnr = _val(r.stdout, "Global PSNR:")
    assert cr > 1.0 and psnr > 40.0
    assert _val(r.stdout, "Auto query final ROI max abs diff:") == 0.0
    assert "Query [3,40] x [4,30] x [5,20] (box)" in r.stdout
    dec = np.fromfile(out, dtype=dt).reshape(f.shape)
    bound = 1e-3 * float(f.max() - f.min()) if mode else eb * 8.2
    assert np.abs(dec.astype(np.float64) - f).max() <= bound * (1 + 1e-9)
    lib_dec = ctx.decompress(np.fromfile(stream, dtype=np.uint8))
    assert np.array_equal(lib_dec.view(np.uint8), dec.view(np.uint8))
