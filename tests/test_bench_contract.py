"""bench.py's contract on the CPU: the kernel names it reports match the
committed rocprofv3 summaries it takes the HBM traffic from, and a request for
more than one GPU ends with a message before anything touches a GPU."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402  (module level imports the standard library only)

I8_DEFAULT = {"kernel": 2, "staging": 2, "stage": 32, "mfma": 0}
F16_DEFAULT = {"kernel": 2, "staging": 2, "stage": 16, "mfma": 0}


def _profile(name):
    return open(os.path.join(ROOT, "profiles", name)).read()


def test_headline_traffic_comes_from_the_committed_profile():
    kname = bench.kernel_name(I8_DEFAULT)
    assert kname == "xg_xengine_rgd<32, 8>"
    text = _profile("xg_xengine_rgd_int8_ns512_nf64_nt8192.md")
    want = int(re.search(r"^traffic_bytes_per_launch: (\d+)$", text, re.M).group(1))
    assert bench.profiled_traffic(kname) == want
    # the traffic is the scaled FETCH_SIZE + WRITE_SIZE: within 2x of the
    # 1.343e9 algorithmic bytes of a launch (input + register-tile output)
    # (matLength = NF x 256*257/2 cells x 4 quadrants x 4 products; re, im int32)
    algorithmic = 512 * 64 * 8192 * 4 + 64 * (256 * 257 // 2) * 16 * 2 * 4
    assert algorithmic == 1_343_225_856
    assert algorithmic <= want <= 2 * algorithmic
    # a summary of another kernel is not taken for this one
    assert bench.profiled_traffic("xg_xengine_rgd<64, 4>") is None


def test_kernel_names_match_the_profiles():
    assert bench.kernel_name({**I8_DEFAULT, "stage": 64}) == "xg_xengine_rgd<64, 4>"
    assert bench.kernel_name({**I8_DEFAULT, "stage": 64, "mfma": 2}) == "xg_xengine_rgd16<4>"
    assert bench.kernel_name({**I8_DEFAULT, "kernel": 3}) == "xg_xengine_kg"
    f16 = bench.f16_kernel_name(F16_DEFAULT)
    assert f16 == "xg_xengine_rgdh<16, 8>"
    assert f"`{f16}`" in _profile("xg_xengine_rgdh_fp16_ns256_nf128_nt2048.md").splitlines()[0]
    assert bench.f16_kernel_name({**F16_DEFAULT, "stage": 32}) == "xg_xengine_rgdh<32, 4>"


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
