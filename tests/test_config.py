"""The YASF configuration reader (yasfpy/config.py:46-480) on the reference's
graphite example (tests/golden/config_graphite: its config.json and cluster
file, copied unchanged) and on a local-table YAML configuration."""

import json
import math
from pathlib import Path

import numpy as np
import numpy.testing as npt
import pytest

from yasf_amd.config import Config, fix_overlap, sphere_overlap_count

GRAPHITE = Path(__file__).resolve().parent / "golden" / "config_graphite"


def _brute_overlaps(pos, r):
    n = r.size
    c = np.zeros(n)
    for i in range(n):
        for j in range(i + 1, n):
            if np.sum((pos[i] - pos[j]) ** 2) < (r[i] + r[j]) ** 2:
                c[i] += 1
                c[j] += 1
    return c


def test_graphite_example_geometry_and_units():
    cfg = Config(str(GRAPHITE / "config.json"), preprocess=False, seed=0)
    raw = np.loadtxt(GRAPHITE / "N_256_FD_20.dat")
    # wavelengths: arange(1, 3.001, 0.5) in micrometres
    npt.assert_allclose(cfg.wavelength, [1.0, 1.5, 2.0, 2.5, 3.0])
    assert cfg.wavelength_scale == 1e-6 and cfg.particles_scale == 1e-9
    assert cfg.cluster_dimensional_scale == 1.0
    # positions and radii from nm to the wavelength unit (um): x 1e-9 / 1e-6
    assert cfg.spheres.shape == (256, 5)
    npt.assert_allclose(cfg.spheres[:, :3], raw[:, :3] * 1e-3, rtol=1e-15)
    # overlapping spheres were shrunk by powers of 0.95, the rest kept
    ratio = cfg.spheres[:, 3] / (raw[:, 3] * 1e-3)
    k = np.log(ratio) / np.log(0.95)
    npt.assert_allclose(k, np.round(k), atol=1e-9)
    assert np.all(ratio <= 1.0 + 1e-15)
    assert not np.any(_brute_overlaps(cfg.spheres[:, :3], cfg.spheres[:, 3]) > 0)
    # one material, uniform distribution: every sphere gets material 0
    assert np.all(cfg.spheres[:, 4] == 0)
    assert cfg.material == [{"url": json.loads((GRAPHITE / "config.json").read_text())
                             ["particles"]["material"][0]["url"]}]
    assert cfg.output_filename.name == "config_N_256_FD_20_dimscale1p00.bz2"
    assert cfg.config["optics"] is True


def test_graphite_example_refuses_to_fetch_urls():
    with pytest.raises(RuntimeError, match="network"):
        Config(str(GRAPHITE / "config.json"))  # preprocess=True needs the URL materials


def test_fix_overlap_matches_brute_force_counting():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 10, size=(40, 3))
    r = rng.uniform(0.5, 1.6, size=40)
    npt.assert_array_equal(sphere_overlap_count(pos, r), _brute_overlaps(pos, r))
    r0 = r.copy()
    fixed = fix_overlap(pos, r)
    assert not np.any(_brute_overlaps(pos, fixed) > 0)
    untouched = _brute_overlaps(pos, r0) == 0
    npt.assert_array_equal(fixed[untouched], r0[untouched])


def test_yaml_config_with_local_tables(tmp_path):
    # material table in micrometres, wavelengths in nanometres from a file,
    # constant medium, dimensional scale, points grid, distribution over two
    # materials given as a fifth column
    (tmp_path / "mat.txt").write_text("# wl[um] n k\n0.4 1.40 0.010\n0.6 1.60 0.030\n0.8 1.70 0.050\n")
    (tmp_path / "lam.csv").write_text("450\n700\n")
    spheres = np.array([[0.0, 0.0, 0.0, 50.0, 0], [300.0, 0.0, 0.0, 60.0, 1], [0.0, 400.0, 0.0, 55.0, 1]])
    np.savetxt(tmp_path / "c.dat", spheres)
    (tmp_path / "run.yaml").write_text(
        "particles:\n"
        "  geometry: {file: c.dat, scale: 1.0e-9, dimensional_scale: 2.0}\n"
        "  material:\n"
        "    - {path: mat.txt, scale: 1.0e-6}\n"
        "    - {constant: '2.0+0.1j'}\n"
        "initial_field: {beam_width: 0, focal_point: [0, 0, 0], polar_angle: 0, azimuthal_angle: 0,"
        " polarization: TE}\n"
        "parameters:\n"
        "  wavelength: {path: lam.csv, scale: 1.0e-9}\n"
        "  medium: 1.33\n"
        "solver: {type: gmres, tolerance: 1.0e-6, max_iter: 100, restart: 50}\n"
        "numerics: {lmax: 3, sampling_points: [10, 19], particle_distance_resolution: 1, gpu: true}\n"
        "points: {x: {start: 0, stop: 3, step: 1}, y: [0, 5], z: 7}\n"
        "output: {folder: out, extension: json, filename: result}\n")
    cfg = Config(str(tmp_path / "run.yaml"))
    npt.assert_allclose(cfg.wavelength, [450.0, 700.0])
    # dimensional scale 2, then nm -> nm (1e-9 / 1e-9)
    npt.assert_allclose(cfg.spheres[:, :4], spheres[:, :4] * 2.0)
    npt.assert_array_equal(cfg.spheres[:, 4], spheres[:, 4])
    # 450 nm = 0.45 um: a quarter of the way from 0.4 to 0.6
    n450 = 1.40 + 0.25 * 0.20 + 1j * (0.010 + 0.25 * 0.020)
    n700 = 1.60 + 0.5 * 0.10 + 1j * (0.030 + 0.5 * 0.020)
    npt.assert_allclose(cfg.refractive_index_interpolated[0], [n450, n700], rtol=1e-14)
    npt.assert_allclose(cfg.refractive_index_interpolated[1], [2.0 + 0.1j] * 2)
    npt.assert_allclose(cfg.medium_refractive_index, [1.33, 1.33])
    p = cfg.particles(1)
    npt.assert_allclose(p.refractive_index, [n700, 2.0 + 0.1j, 2.0 + 0.1j], rtol=1e-14)
    npt.assert_allclose(p.r, [100.0, 120.0, 110.0])
    assert cfg.config["points_shape"] == (3, 2, 1)
    assert cfg.config["points"]["x"].size == 6 and math.isclose(cfg.config["points"]["z"][0], 7.0)
    assert cfg.output_filename == Path.cwd() / "out" / "result.json"


def test_config_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        Config(str(tmp_path / "run.txt"))
    with pytest.raises(TypeError):
        Config(Path("x.json"))
    (tmp_path / "bad.json").write_text(json.dumps({"particles": {"geometry": {}}}))
    with pytest.raises(FileNotFoundError, match="cluster file"):
        Config(str(tmp_path / "bad.json"))
