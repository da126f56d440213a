"""The recon program end to end on the GPU (Program/Main.cpp:98-197 through
the C++ API: ReconData -> GridLut -> ImageFilter), SOS and PILS, against the
oracle's restatement of the same pipeline.

Bar: relative L2 <= 1e-5 on the written |image| (the oracle's FFT is FP64)."""
import os
import subprocess

import numpy as np
import pytest

from cli_files import write_acquisition
from fastrecon_amd import synthetic as syn

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "fastrecon_amd", "fastrecon")
TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


CASES = [
    # dim, spokes, readout, coils, image, pils
    (2, 402, 512, 8, 128, False),
    (2, 302, 384, 4, 192, False),      # 384^2 grid: the radix-6,8,8 cropped transforms behind ImageFilter::SOS
    (2, 201, 256, 4, 64, True),
    (3, 800, 64, 4, 32, False),
    (3, 800, 64, 3, 32, True),
]


@pytest.mark.parametrize("dim,spokes,readout,coils,image,pils", CASES)
def test_program_matches_oracle(orc, tmp_path, dim, spokes, readout, coils, image, pils):
    traj, dcf = syn.radial2d(spokes, readout) if dim == 2 else syn.kooshball3d(spokes, readout)
    M = traj.shape[0]
    data = syn.coil_data(coils, M, 33)
    rc = (image, image, image if dim == 3 else 1)
    ini = write_acquisition(str(tmp_path), traj, dcf, data, readout, spokes, rc)
    args = [CLI, ini] + (["--pils"] if pils else [])
    print(f"[cli] dim={dim} M={M} C={coils} image={image} pils={pils}", flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    img = np.fromfile(os.path.join(str(tmp_path), "image.raw"), np.float32)
    assert img.size == rc[0] * rc[1] * rc[2]

    N = syn.grid_size(image, 2.0)
    lut = orc.kb_lut(4.0, 2.0, 512)[0]
    tn = orc.normalize_traj(traj, N)[0]
    if pils:
        ref = orc.pils_f32(tn, dcf, data, N, 4.0, lut, rc, nthreads=8)
    else:
        ref = orc.recon_f32(tn, dcf, data, N, 4.0, lut, rc, nthreads=8)
    e = rel_l2(img, ref)
    print(f"[cli] image rel-L2 vs oracle = {e:.2e}", flush=True)
    assert e <= TOL
