"""YASF configuration reader (mirror of yasfpy/config.py:46-480, ``Config``).

Reads the reference's JSON or YAML run configuration: the particle geometry
section (cluster file, delimiter, ``scale``, ``dimensional_scale``), materials
and their ``distribution``, ``initial_field``, ``parameters`` (wavelength as a
list, an arange dict or a file; ``scale``; medium), ``solver``, ``numerics``,
``optics``, ``points`` and ``output``.  Units follow the reference: sphere
positions and radii are multiplied by ``dimensional_scale`` and then by
``particles scale / wavelength scale`` (config.py:250-271), so that they are in
the unit of the wavelengths.

Refractive-index data: the reference fetches ``url`` materials through
refidxdb (config.py:166-196).  Nothing here goes to the network: a ``url``
material or medium is kept as a reference and raises ``RuntimeError`` when its
values are needed (``preprocess=True`` or ``interpolate()``).  A ``path``
material is read from a local three-column table (wavelength, n, k; text or
CSV, ``#`` comments) in the material's length unit ``scale``; a ``constant``
key (n + ik, an extension) gives a wavelength-independent index.  A number as
medium is used as is, as in the reference.
"""

from __future__ import annotations

import json
import logging
from numbers import Number
from pathlib import Path

import numpy as np

from yasf_amd.aggregates import read_cluster

OVERLAP_MULTIPLICATOR = 0.95  # config.py:22


def sphere_overlap_count(positions: np.ndarray, radii: np.ndarray) -> np.ndarray:
    """Per sphere, the number of spheres it overlaps (|r_i - r_j| < r_i + r_j),
    as sphere_overlap_check (config.py:483-504) counts them."""
    positions = np.asarray(positions, dtype=float)
    radii = np.asarray(radii, dtype=float)
    counter = np.zeros(radii.size)
    for i in range(radii.size - 1):
        d2 = np.sum((positions[i + 1:] - positions[i]) ** 2, axis=1)
        hit = d2 < (radii[i + 1:] + radii[i]) ** 2
        counter[i] += np.count_nonzero(hit)
        counter[i + 1:] += hit
    return counter


def fix_overlap(positions: np.ndarray, radii: np.ndarray, max_iter: int = 100) -> np.ndarray:
    """Shrinks, up to max_iter times, the radii of the spheres with the most
    overlaps by OVERLAP_MULTIPLICATOR until none overlap (config.py:508-531).
    Modifies and returns radii."""
    for _ in range(max_iter):
        counter = sphere_overlap_count(positions, radii)
        if not np.any(counter > 0):
            break
        radii[counter == counter.max()] *= OVERLAP_MULTIPLICATOR
    return radii


def _arange_or_list(value, what: str) -> np.ndarray:
    if isinstance(value, Number):
        return np.array([value], dtype=float)
    if isinstance(value, list):
        return np.array(value, dtype=float)
    if isinstance(value, dict):
        return np.arange(value["start"], value["stop"], value["step"])
    raise ValueError(f"{what}: numbers, lists of numbers and (start, stop, step) dicts are permitted")


class Config:
    """Run configuration of one YASF simulation (yasfpy/config.py:46-139)."""

    def __init__(self, path_config: str = "", preprocess: bool = True, path_cluster: str = "",
                 cluster_scale: float | None = None, cluster_dimensional_scale: float | None = None, *,
                 quiet: bool = False, seed: int | None = None):
        if not isinstance(path_config, str):
            raise TypeError("The config file path needs to be a string!")
        self.path_config = Path(path_config)
        self.file_type = self.path_config.suffix
        if self.file_type == ".json":
            with open(self.path_config) as fh:
                self.config = json.load(fh)
        elif self.file_type in (".yaml", ".yml"):
            import yaml

            with open(self.path_config) as fh:
                self.config = yaml.safe_load(fh)
        else:
            raise FileNotFoundError("The provided config file needs to be a json or yaml file!")
        if self.config is None:
            raise FileNotFoundError(f"Could not read config file {path_config}.")
        geometry = self.config.get("particles", {}).get("geometry", {})
        if path_cluster != "":
            self.path_cluster = Path(path_cluster)
        elif "file" in geometry:
            self.path_cluster = self.path_config.parent / geometry["file"]
        else:
            raise FileNotFoundError("The cluster file needs to be provided either in the config file "
                                    "or as an argument.")
        self.particles_scale = float(geometry.get("scale", 1.0) if cluster_scale is None else cluster_scale)
        self.cluster_dimensional_scale = float(geometry.get("dimensional_scale", 1.0)
                                               if cluster_dimensional_scale is None
                                               else cluster_dimensional_scale)
        self.log = logging.getLogger(self.__class__.__module__)
        if quiet:
            self.log.setLevel(logging.WARNING)
        self._rng = np.random.default_rng(seed)
        self.refractive_index_interpolated = None
        self.medium_refractive_index = None
        self._read()
        self._output()
        if preprocess:
            self.interpolate()

    # ------------------------------------------------------------ sections

    def _read(self):
        cfg = self.config
        wl = cfg["parameters"]["wavelength"]
        if "path" in wl:
            path = Path(wl["path"])
            self.wavelength = np.atleast_1d(np.loadtxt(path if path.is_absolute()
                                                       else self.path_config.parent / path))
        elif isinstance(wl.get("data"), (list, dict)):
            self.wavelength = _arange_or_list(wl["data"], "wavelength")
        else:
            raise ValueError("Please provide the wavelength data as an array, or the (start, stop, step) "
                             "numpy.arange parameters.")
        self.wavelength_scale = float(wl.get("scale", 1))

        self.material = []
        for mat in cfg["particles"]["material"]:
            if "url" in mat:
                self.material.append({"url": mat["url"]})
            elif "path" in mat:
                path = Path(mat["path"])
                self.material.append({"path": path if path.is_absolute() else self.path_config.parent / path})
            elif "constant" in mat:
                self.material.append({"constant": complex(mat["constant"])})
            else:
                raise ValueError("Material needs to have a url or path key.")
        self.material_scale = [float(mat.get("scale", 1.0)) for mat in cfg["particles"]["material"]]
        self.medium = cfg["parameters"]["medium"]

        spheres, self.particle_metadata = read_cluster(self.path_cluster, geometry_delimiter(cfg))
        spheres = np.array(spheres, dtype=float)
        if spheres.shape[1] == 4:
            dist = cfg["particles"].get("distribution")
            if dist is None:
                self.log.warning("Implying a uniform distribution of all materials among the particles")
                dist = np.ones(len(self.material)) / len(self.material)
                cfg["particles"]["distribution"] = dist
            dist = np.asarray(dist, dtype=float)
            spheres = np.hstack((spheres, self._rng.choice(dist.size, size=(spheres.shape[0], 1), p=dist)))
        elif spheres.shape[1] > 5:
            self.log.warning("More than 5 columns have been provided. Everything after the 5th will be ignored!")
            spheres = spheres[:, :5]
        spheres[:, :4] *= self.cluster_dimensional_scale
        spheres[:, 3] = fix_overlap(spheres[:, :3], spheres[:, 3])
        factor = self.particles_scale / self.wavelength_scale
        meta = self.particle_metadata
        if meta and "aggregate_properties" in meta and "radius_of_gyration" in meta["aggregate_properties"]:
            meta["aggregate_properties"]["radius_of_gyration"] *= factor
        spheres[:, :4] *= factor
        self.spheres = spheres

        optics = cfg.get("optics", True)
        cfg["optics"] = optics if isinstance(optics, bool) else True
        if "points" in cfg:
            pts = {"x": np.array([0.0]), "y": np.array([0.0]), "z": np.array([0.0])}
            for key, value in cfg["points"].items():
                pts[key] = _arange_or_list(value, f"points {key}")
            x, y, z = np.meshgrid(pts["x"], pts["y"], pts["z"], indexing="ij")
            cfg["points"] = {"x": x.ravel(), "y": y.ravel(), "z": z.ravel()}
            cfg["points_shape"] = x.shape

    def _output(self):
        out = self.config.get("output", {})
        if isinstance(out, str):
            out = {"filename": out}
        folder = Path(out.get("folder", "."))
        if not folder.is_absolute():
            folder = Path.cwd() / folder
        ext = out.get("extension", "pbz2")
        name = f"{self.path_config.stem}_{self.path_cluster.stem}_dimscale{self.cluster_dimensional_scale:.2f}"
        name = out.get("filename", name.replace(".", "p"))
        # the reference creates the folder here (config.py:362-371); it is
        # only named, and created by whoever writes the output
        self.output_filename = folder / f"{name}.{ext}"

    # ------------------------------------------------------------ refractive indices

    def _table(self, entry, scale: float, what: str):
        if "url" in entry:
            raise RuntimeError(f"{what}: refractive-index data from {entry['url']} would have to be fetched; "
                               "this reader does not use the network -- give a local 'path' table "
                               "(wavelength, n, k) or a 'constant'")
        if "constant" in entry:
            return None, entry["constant"]
        table = np.loadtxt(entry["path"], delimiter="," if Path(entry["path"]).suffix == ".csv" else None,
                           comments="#", ndmin=2)
        if table.shape[1] < 3:
            raise ValueError(f"{what}: the table needs wavelength, n and k columns")
        return table, None

    def _interp(self, entry, scale: float, what: str) -> np.ndarray:
        table, const = self._table(entry, scale, what)
        if table is None:
            return np.full(self.wavelength.size, const, dtype=complex)
        w = table[:, 0] * scale
        nk = table[:, 1] + 1j * table[:, 2]
        order = np.argsort(w)
        w, nk = w[order], nk[order]
        target = self.wavelength * self.wavelength_scale
        return np.interp(target, w, nk.real) + 1j * np.interp(target, w, nk.imag)

    def interpolate(self):
        """Refractive index of every material (materials x wavelengths) and of the
        medium at the configured wavelengths (config.py:394-431)."""
        self.refractive_index_interpolated = np.array(
            [self._interp(m, s, f"material {i}") for i, (m, s) in enumerate(zip(self.material,
                                                                                 self.material_scale))],
            dtype=complex)
        med = self.medium
        if isinstance(med, Number):
            self.medium_refractive_index = np.full(self.wavelength.size, float(med), dtype=complex)
        elif isinstance(med, dict):
            self.medium_refractive_index = self._interp(med, float(med.get("scale", 1.0)), "medium")
        else:
            raise ValueError("medium: a number or a dict with url/path/constant")
        return self.refractive_index_interpolated

    # ------------------------------------------------------------ model objects

    def particles(self, wavelength_index: int = 0):
        """Particles at one configured wavelength: positions and radii in the
        wavelength unit, each sphere's index from its material's table."""
        from yasf_amd.model import Particles

        if self.refractive_index_interpolated is None:
            self.interpolate()
        material = self.spheres[:, 4].astype(int)
        n = self.refractive_index_interpolated[material, wavelength_index]
        return Particles(self.spheres[:, :3], self.spheres[:, 3], n)


def geometry_delimiter(cfg: dict) -> str | None:
    return cfg["particles"]["geometry"].get("delimiter")
