"""The C++ API on the GPU (tests/cpp/gridlut_api_check.cpp, built by
fastrecon_amd/build.py): device plans shared between GridLut instances on the
same trajectory, as the reference shares its gridding map between instances
and frames (cuGridLut.cu:156-159, cuGridLut.h:50-56, GridLut.h:24-26); a
second trajectory gets its own plan; a shared plan grids bit-identically to a
fresh one; griddingChannel over all channels costs one gridding pass
(GridLut.cpp:53-78) and sees data updates."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "gridlut_api_check")


def test_plan_sharing_and_channel_cache():
    assert os.path.exists(EXE), "tests/cpp/gridlut_api_check is not built (fastrecon_amd/build.py)"
    r = subprocess.run([EXE, "4000", "128", "4", "64"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    res = json.loads(line)
    print(f"[c++ api] {res}", flush=True)
    assert res["shared"] and res["other_traj_own_plan"] and res["plans_held"] == 2
    assert res["rebuilt_after_release"] and res["bitwise_shared_vs_fresh"]
    assert res["channel_loop_shares_plan"] and res["channel_loop_passes"] == 1 and res["channels_equal_execute"]
    assert res["passes_after_update"] == 2 and res["update_seen"]
    # the shared plan costs a trajectory hash, not a plan build
    assert res["plan_shared_ms"] < res["plan_first_ms"]
