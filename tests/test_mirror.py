"""The C++ mirror of the reference interfaces (par-cd_amd/host/parcd.hpp)
used the way a reference user writes a driver: tests/cpp/mirror_driver.cpp
runs the main.cpp:226-266 loop through Graph, Partition, the rule classes,
the refinement classes and the framework functions, and declares a rule of
its own with the five methods of rules.hpp:9-19.

On the CPU the driver must compile against the mirror (a rule written to
rules.hpp included).  On the GPU its labels must equal the oracle's serial
replay, and the device must refuse the rule it cannot run."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import HERE, ROOT, golden, load

SRC = os.path.join(HERE, "cpp", "mirror_driver.cpp")
PKG = os.path.join(ROOT, "par-cd_amd")


@pytest.fixture(scope="module")
def driver():
    """tests/cpp/mirror_driver, built by the build step (tests/cpp/Makefile)."""
    exe = os.path.join(HERE, "cpp", "mirror_driver")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/mirror_driver not built (run the build step)")
    return exe


def test_reference_style_driver_and_rule_compile(tmp_path):
    if not os.path.exists(os.path.join(PKG, "libparcd.so")):
        pytest.skip("libparcd.so not built")
    exe = str(tmp_path / "mirror_driver")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-Wno-unused-parameter",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc"), "-I" + os.path.join(PKG, "host"),
           SRC, "-L" + PKG, "-lparcd", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.access(exe, os.X_OK)


def _graph(tmp_path, name):
    d = load(golden("graph_%s.npz" % name)[0])
    p = tmp_path / (name + ".pcd")
    p.write_bytes(d["pcd"].tobytes())
    return str(p), O.parse_pcd(d["pcd"].tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("name,iters,metric,strategy", [
    ("sbm200", 6, "mod", 0),
    ("rmat9", 5, "cpm-0.05", 0),
    ("sbm210", 4, "mod", 1),
    ("gnp150", 3, "cpm-0.02", 1),
])
def test_mirror_driver_equals_replay(driver, tmp_path, name, iters, metric, strategy):
    path, (n, m, ind, adj, mp) = _graph(tmp_path, name)
    r = subprocess.run([driver, path, str(iters), metric, str(strategy)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr[-800:]
    got = np.array([int(x) for x in r.stdout.split()], np.int32)
    kind, lam = (O.MOD, 0.0) if metric == "mod" else (O.CPM, float(metric[4:]))
    # main.cpp:226-266 with -n iters: the labels compressed after the last step
    want, _, _, _, _ = O.run(n, m, ind, adj, kind=kind, lam=lam, strategy=strategy, max_iter=iters, mode=1, seed=0,
                             tie=O.TIE_LOWEST)
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_mirror_refuses_rule_without_device_form(driver, tmp_path):
    path, _ = _graph(tmp_path, "sbm200")
    r = subprocess.run([driver, path, "2", "mod", "0", "custom"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, (r.stdout[-300:], r.stderr[-300:])
    assert r.stdout.strip() == "refused %d" % -1
    # the host half of the rule ran before the refusal
    assert "half modularity" in r.stderr
