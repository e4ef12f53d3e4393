"""The fine_registration and normal_estimation tools end to end: PLY in, GPU
ICP / normals, PLY + stdout out, compared with the Python binding and the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from conftest import ROOT, print_case
from lowcost3dreconstruction_amd import synth

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "lowcost3dreconstruction_amd", "bin")


def write_ascii_ply(path, xyz, nrm=None, rgb=None):
    """The capture tools' ASCII layout: x y z nx ny nz red green blue
    (capture/depth_capture/depth_capture.cpp:281-306)."""
    n = len(xyz)
    nrm = np.zeros_like(xyz) if nrm is None else nrm
    rgb = np.full((n, 3), 128, np.uint8) if rgb is None else rgb
    with open(path, "w") as f:
        f.write("ply\nformat ascii 1.0\nelement vertex %d\n" % n)
        for p in ("x", "y", "z", "nx", "ny", "nz"):
            f.write("property float %s\n" % p)
        f.write("property uchar red\nproperty uchar green\nproperty uchar blue\nend_header\n")
        for i in range(n):
            f.write("%.9g %.9g %.9g %.9g %.9g %.9g %d %d %d\n" % (*xyz[i], *nrm[i], *rgb[i]))


def read_binary_ply(path):
    with open(path, "rb") as f:
        data = f.read()
    head, body = data.split(b"end_header\n", 1)
    lines = head.decode().splitlines()
    n = int([l for l in lines if l.startswith("element vertex")][0].split()[-1])
    rec = struct.Struct("<3f3B4f")
    out = np.empty((n, 7), np.float64)
    rgb = np.empty((n, 3), np.uint8)
    for i in range(n):
        v = rec.unpack_from(body, i * rec.size)
        out[i] = [v[0], v[1], v[2], v[6], v[7], v[8], v[9]]
        rgb[i] = v[3:6]
    assert "property float normal_x" in lines and "element camera 1" in lines
    return out, rgb


def test_fine_registration_cli(tmp_path):
    p = synth.pair_ellipsoids(9, n_src=3000, n_tgt=3000)
    src, tgt = tmp_path / "src.ply", tmp_path / "tgt.ply"
    rgb = np.random.default_rng(0).integers(0, 255, size=(3000, 3)).astype(np.uint8)
    write_ascii_ply(src, p.src, rgb=rgb)
    write_ascii_ply(tgt, p.tgt)
    out, acc, mat = tmp_path / "out.ply", tmp_path / "acc.ply", tmp_path / "T.txt"
    print_case(test="cli_fine_registration")
    r = subprocess.run([os.path.join(BIN, "fine_registration"), "-i", str(src), "-t", str(tgt),
                        "-o", str(out), "-a", str(acc), "--distance_threshold", "0.05",
                        "--matrix_out", str(mat)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == f"Loaded 3000 data points from {src}"
    assert lines[1] == f"Loaded 3000 data points from {tgt}"
    assert lines[2] in ("Has converged: True", "Has converged: False")
    assert lines[3].startswith("Score: ")
    printed = np.array([[float(x) for x in l.split()] for l in lines[4:8]])
    T = np.loadtxt(mat)
    ref = oracle.icp(p.src, p.tgt, max_corr_dist=0.05, max_iterations=50)
    er, et = synth.transform_error(T, ref["T"])
    assert er < 1e-6 and et < 1e-7
    np.testing.assert_allclose(printed, T, atol=2e-6)  # 6 significant digits, like Eigen
    assert abs(float(lines[3].split()[1]) - ref["fitness"]) <= 1e-5 * ref["fitness"]
    reg, reg_rgb = read_binary_ply(out)
    np.testing.assert_array_equal(reg[:, :3].astype(np.float32), oracle.transform_points(p.src, T))
    np.testing.assert_array_equal(reg_rgb, rgb)  # colour survives the round trip
    accd, _ = read_binary_ply(acc)
    assert len(accd) == 6000
    np.testing.assert_array_equal(accd[3000:, :3].astype(np.float32), p.tgt)


def test_fine_registration_cli_errors(tmp_path):
    exe = os.path.join(BIN, "fine_registration")
    r = subprocess.run([exe, "-i", "a.ply"], capture_output=True, text=True)
    assert r.returncode == 255 and "Correct mode of use" in r.stderr
    r = subprocess.run([exe, "-i", "a", "-t", "b", "-o", "c", "--max_iterations", "0"],
                       capture_output=True, text=True)
    assert r.returncode == 255 and "max_iterations needs to be greater than zero." in r.stderr
    r = subprocess.run([exe, "-i", str(tmp_path / "missing.ply"), "-t", "b", "-o", "c"],
                       capture_output=True, text=True)
    assert r.returncode == 255 and "Couldn't load input cloud file" in r.stderr


@pytest.mark.parametrize("flags,flip", [([], False), (["-r"], True), (["-c"], True),
                                        (["-c", "-r"], False), (["-z"], False)])
def test_normal_estimation_cli(tmp_path, lcr, flags, flip):
    p = synth.pair_depth(2, scale=1.0)
    xyz = p.tgt[:4000]
    src, out = tmp_path / "in.ply", tmp_path / "out.ply"
    write_ascii_ply(src, xyz, nrm=np.ones_like(xyz))
    print_case(test="cli_normals", flags=" ".join(flags))
    r = subprocess.run([os.path.join(BIN, "normal_estimation"), "-i", str(src), "-o", str(out),
                        "-n", "20"] + flags, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, _ = read_binary_ply(out)
    vp = xyz.astype(np.float64).mean(0).astype(np.float32) if "-c" in flags else (0, 0, 0)
    rn, rc = oracle.normals(xyz, 20, vp)
    if flip:
        rn = -rn
    dots = np.sum(got[:, 3:6] * rn, axis=1)
    assert np.mean(dots > 1 - 1e-6) > 0.999
