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


def test_transform_points_matches_oracle(lcr):
    rng = np.random.default_rng(2)
    xyz = rng.normal(size=(5000, 3)).astype(np.float32)
    xyz[7] = np.nan
    nrm = rng.normal(size=(5000, 3)).astype(np.float32)
    T = synth.rigid(synth.random_rotation(rng), rng.normal(size=3))
    print_case(test="transform_points")
    out, out_n = lcr.transform_points(xyz, T, nrm)
    ref = oracle.transform_points(xyz, T)
    np.testing.assert_array_equal(out[np.isfinite(ref)], ref[np.isfinite(ref)])
    assert np.all(np.isnan(out[7]))
    np.testing.assert_allclose(out_n, nrm.astype(np.float64) @ T[:3, :3].T, rtol=0, atol=2e-6)


def test_align_cycle_cli_chains_poses(lcr, tmp_path):
    """alignment.sh's naming (0.ply .. N-1.ply, top.ply, bottom.ply): one batched
    run, every pair equal to the Cycle binding from the same seeds (the turntable
    prior; centroid_align's transform for top and bottom), poses chained into
    capture 0's frame with the closing pair's residual spread over the loop."""
    centre = (0.0, 0.0, -0.8)
    cyc = synth.turntable_cycle(8, n_captures=4, yaw_step_deg=10.0, scale=1.5, centre=centre,
                                top_bottom=True, closing_pair=True)
    names = ["0", "1", "2", "3", "top", "bottom"]
    for name, c in zip(names, cyc.clouds):
        write_ascii_ply(tmp_path / f"{name}.ply", c)
    out = tmp_path / "out"
    out.mkdir()
    print_case(test="cli_align_cycle")
    # the synthetic top / bottom views are tilted by +-45 deg about the camera x axis
    r = subprocess.run([os.path.join(BIN, "align_cycle"), "-d", str(tmp_path), "-n", "4", "-o",
                        str(out), "--top_bottom", "--closing", "--point_to_plane",
                        "--distance_threshold", "0.01", "--max_iterations", "60", "--yaw_step",
                        "10", "--centre", "0,0,-0.8", "--top_seed", "-45,0,0,0", "--bottom_seed",
                        "45,0,0,0"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("Has converged:") == len(cyc.pairs)
    prior = synth.turntable_prior(10.0, centre)
    guesses = np.stack([prior] * 4 + [lcr.centroid_align(cyc.clouds[4], cyc.clouds[0], -45, 0, 0, 0),
                                      lcr.centroid_align(cyc.clouds[5], cyc.clouds[0], 45, 0, 0, 0)])
    ref = lcr.Cycle(cyc.clouds, cyc.pairs, guesses=guesses, point_to_plane=True,
                    max_corr_dist=0.01, max_iterations=60, normals_k=30)
    ref.run()
    T = ref.results()["T"]
    chain = lcr.chain_poses(T[:3], T[3])
    pose = {str(i): chain[i] for i in range(4)}
    pose["top"], pose["bottom"] = T[4], T[5]
    for name, c in zip(names, cyc.clouds):
        M = np.loadtxt(out / f"{name}.txt")
        np.testing.assert_allclose(M, pose[name], rtol=0, atol=1e-12)
        got, _ = read_binary_ply(out / f"{name}.ply")
        exp = oracle.transform_points(c, M)
        np.testing.assert_array_equal(got[:, :3].astype(np.float32), exp.astype(np.float32))
    # the turntable captures against the true motion: capture j in capture 0's frame
    # is poses[0] poses[j]^-1 (top and bottom: the seed test below; centroids of
    # differently cropped partial views are too far apart for a 1 cm radius)
    for i, name in enumerate(names[:4]):
        truth = cyc.poses[0] @ np.linalg.inv(cyc.poses[i])
        er, et = synth.transform_error(pose[name], truth)
        assert er < np.deg2rad(0.5) and et < 3e-3, (name, np.rad2deg(er), et)


def test_align_cycle_cli_top_bottom_seed_recovers_true_pose(tmp_path):
    """Top and bottom views held the way alignment.sh:117-118 expects (turned by the
    script's roll / yaw about their centroid): centroid_align seeds with the
    script's angles (-90,0,180 and 90,0,0; elevation 0 here, as every view samples
    the whole surface) bring them within reach of ICP, which recovers the true
    pose, not just the binding's.  Moved onto the centroid without the rotation,
    ICP cannot."""
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(21)
    shape = synth.make_shape(rng)

    def sample():
        return synth.sample_surface(shape, 20000, rng) + rng.normal(scale=1e-3, size=(20000, 3))

    cap0 = sample()
    c = cap0.mean(0)
    yaw = synth.turntable_prior(10.0, c)
    views = {"0": cap0, "1": synth.apply(np.linalg.inv(yaw), sample())}
    truth = {"0": np.eye(4), "1": yaw}
    for name, (roll, ryaw), d in (("top", (-90.0, 180.0), (2.0, 0.003)),
                                  ("bottom", (90.0, 0.0), (-2.5, -0.002))):
        R = Rotation.from_euler("XYZ", [roll, 0.0, ryaw], degrees=True).as_matrix()
        delta = synth.rigid(synth.rot_axis_angle([1, 1, 0], np.deg2rad(d[0])), [d[1], 0, d[1]])
        C = synth.rigid(np.eye(3), c)
        T_gt = C @ delta @ synth.rigid(R, np.zeros(3)) @ np.linalg.inv(C)
        views[name] = synth.apply(np.linalg.inv(T_gt), sample())
        truth[name] = T_gt
    for name, v in views.items():
        write_ascii_ply(tmp_path / f"{name}.ply", v.astype(np.float32))
    print_case(test="cli_top_bottom_seed")
    for seeded in (True, False):
        out = tmp_path / ("out_seeded" if seeded else "out_identity")
        out.mkdir()
        extra = (["--top_seed", "-90,0,180,0", "--bottom_seed", "90,0,0,0"] if seeded else
                 ["--top_seed", "0,0,0,0", "--bottom_seed", "0,0,0,0"])
        # run to convergence (the CLI's default 1e-3 relative MSE stop leaves a
        # slowly sliding rotation a few tenths of a degree short)
        r = subprocess.run([os.path.join(BIN, "align_cycle"), "-d", str(tmp_path), "-n", "2",
                            "-o", str(out), "--top_bottom", "--distance_threshold", "0.05",
                            "--max_iterations", "300", "--euclidean_fitness_epsilon", "1e-9",
                            "--transformation_epsilon", "1e-14", "--yaw_step", "10", "--centre",
                            ",".join(f"{x:.9g}" for x in c)] + extra,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        for name in ("1", "top", "bottom"):
            er, et = synth.transform_error(np.loadtxt(out / f"{name}.txt"), truth[name])
            print(f"seeded={seeded} {name}: {np.rad2deg(er):.4f} deg {et * 1e3:.4f} mm")
            # ICP's own accuracy on these independent samples of near-symmetric
            # ellipsoids: the turntable pair, started at the exact motion, ends
            # 0.25 deg off; top and bottom must do as well
            if seeded:
                assert er < np.deg2rad(0.5) and et < 2e-3, (name, np.rad2deg(er), et)
            elif name != "1":
                assert er > np.deg2rad(10), (name, np.rad2deg(er))


def test_cloud_downsampling_and_outlier_removal_cli(tmp_path):
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 0.3, size=(4000, 3)).astype(np.float32)
    pts = np.concatenate([pts, rng.uniform(2, 3, size=(10, 3)).astype(np.float32)])
    src = tmp_path / "in.ply"
    write_ascii_ply(src, pts)
    print_case(test="cli_filters")
    out = tmp_path / "down.ply"
    r = subprocess.run([os.path.join(BIN, "cloud_downsampling"), "-i", str(src), "-o", str(out),
                        "-s", "0.02"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Cloud before filtering:" in r.stdout and "Cloud after filtering:" in r.stdout
    got, _ = read_binary_ply(out)
    np.testing.assert_array_equal(got[:, :3].astype(np.float32), oracle.voxel_grid(pts, 0.02)["xyz"])
    out2 = tmp_path / "clean.ply"
    r = subprocess.run([os.path.join(BIN, "outlier_removal"), "-i", str(src), "-o", str(out2),
                        "-n", "20", "-d", "1.0", "-f"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    keep, _, _ = oracle.sor(pts, 20, 1.0)
    kept, _ = read_binary_ply(out2)
    removed, _ = read_binary_ply(tmp_path / "clean_outliers.ply")
    assert len(kept) == keep.sum() and len(removed) == (~keep).sum()
    np.testing.assert_array_equal(kept[:, :3].astype(np.float32), pts[keep])
