"""Structured-box DSS groups (geoflow_amd/csrc/gf_boxdss.hpp) on the CPU.

The gather kernels form each node's GGFX group from the element lattice
(box_members_own) instead of reading the CSR maps.  tests/boxdss_check.cpp
compiles that header for the host together with the product's grid and GGFX
builders (gf_setup.cpp) and checks, node by node over periodic, Dirichlet and
degenerate (one element per direction) boxes at orders 1-7, that the members
equal the GGFX group in ascending order with 64- and 32-bit indices, and that
the reported own slot holds the node itself; at p = 7 it also fills the N = 8
stage kernels' LDS halo with global indices (own values copied, partner
values by Halo8Rim lane by lane) and checks that halo8_order's slots give
every node's group in the same order.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_box_members_match_ggfx_groups(tmp_path):
    exe = str(tmp_path / "boxdss_check")
    cmd = ["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(ROOT, "geoflow_amd", "csrc"), "-I", "/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "boxdss_check.cpp"),
           os.path.join(ROOT, "geoflow_amd", "csrc", "gf_setup.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 20, r.stdout
    # both kinds of N = 8 box occur: the halo order holds, and it does not
    assert "halo_ok=1" in r.stdout and "p=7 ne=2,1,3" in r.stdout
