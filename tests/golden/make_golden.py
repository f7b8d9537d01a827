"""Convert the reference's own test fixtures into small .npz vectors.

Reads the MATLAB/CELES data files the reference's tests hold
(``tests/data/*.mat`` and the CELES FITS file under /root/reference) and writes
plain arrays (inputs and expected outputs only) next to this script.  Run once
in a container where /root/reference exists:

    python tests/golden/make_golden.py

The reference's source is not read or executed; only its data files are.
Layout conversions mirror what the reference's tests do (cited per fixture).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
from scipy.io import loadmat

REF = Path("/root/reference/tests/data")
OUT = Path(__file__).resolve().parent


def _mat(name):
    return loadmat(REF / name)


def make_index():
    d = _mat("multi2single_index_data.mat")  # tests/test_comparison.py:59-67
    np.savez(OUT / "index.npz", lmax=int(d["lmax_msi"][0, 0]),
             inputs=d["inputs_msi"].astype(np.int64),
             outputs=d["outputs_msi"].ravel().astype(np.int64))


def make_legendre():
    d = _mat("legendre_normalized_trigon_data.mat")  # tests/test_comparison.py:70-76
    lmax = int(d["lmax_lnt"][0, 0])
    theta = d["theta_lnt"].ravel()
    cell = d["plm_lnt"]
    plm = np.zeros((lmax + 1, lmax + 1, theta.size))
    for l in range(lmax + 1):
        for m in range(l + 1):
            if cell[l, m].shape:
                plm[l, m, :] = np.asarray(cell[l, m], dtype=float).ravel()
    np.savez(OUT / "legendre.npz", lmax=lmax, theta=theta, plm=plm)


def make_sph_bessel():
    d = _mat("sph_bessel_data.mat")  # tests/test_comparison.py:123-132
    np.savez(OUT / "sph_bessel.npz", l=int(d["l_sb"][0, 0]), z=d["Z_sb"].ravel(),
             j=d["output_sb_nu1"].ravel(), h=d["output_sb_nu3"].ravel())


def make_t_entry():
    d = _mat("t_entry_data.mat")  # tests/test_comparison.py:92-110
    np.savez(OUT / "t_entry.npz", l=int(d["l_te"][0, 0]), k_medium=float(d["kM_te"][0, 0]),
             k_sphere=complex(d["kS_te"][0, 0]), radius=float(d["R_te"][0, 0]),
             scattered_1=complex(d["output_te_1_scattered"].ravel()[0]),
             scattered_2=complex(d["output_te_2_scattered"].ravel()[0]))


def make_spherical_functions():
    d = _mat("spherical_functions_trigon_data.mat")
    arrays = {k: v for k, v in d.items() if not k.startswith("__") and v.dtype != object}
    cells = {k: v for k, v in d.items() if not k.startswith("__") and v.dtype == object}
    n_angles = d["theta_sft"].size  # tests/test_comparison.py:79-89
    for k, v in cells.items():
        lmax = v.shape[0] - 1
        out = np.zeros((lmax + 1, lmax + 1, n_angles))
        for l in range(lmax + 1):
            for m in range(l + 1):
                if v[l, m].shape and v[l, m].size:
                    out[l, m, :] = np.asarray(v[l, m], dtype=float).ravel()
        arrays[k] = out
    np.savez(OUT / "spherical_functions_trigon.npz", **arrays)


def make_transformation():
    d = _mat("transformation_coefficients_data.mat")  # tests/test_comparison.py:157-185
    np.savez(OUT / "transformation_coefficients.npz",
             **{k: np.asarray(v) for k, v in d.items() if not k.startswith("__")})


def make_ab5():
    d = _mat("translation_table_ab_data.mat")  # tests/test_comparison.py:113-120
    np.savez(OUT / "ab5_lmax3.npz", lmax=int(d["lmax_tta"][0, 0]),
             ab5=d["translation_table_tta"]["ab5"][0, 0])


def make_coupling():
    d = _mat("coupling_matrix_multiply_data.mat")  # tests/test_comparison.py:188-228
    lmax = int(d["lmax"][0, 0])
    pos = d["particles_position"].astype(float)
    N = pos.shape[0]
    nmax = 2 * lmax * (lmax + 2)
    # MATLAB [n, s] order -> particle-major [s, n] (tests/test_comparison.py:199-207)
    x = d["input_x"].ravel().reshape(nmax, N).T.ravel()
    wx = d["coupling_matrix_output_Wx"].ravel().reshape(nmax, N).T.ravel()
    np.savez(OUT / "coupling_matvec.npz", lmax=lmax, wavelength=float(d["wavelength"][0, 0]),
             medium_refractive_index=float(d["mediumRefractiveIndex"][0, 0]),
             positions=pos, radii=d["particles_radius"].ravel().astype(float),
             refractive_index=d["particles_refractiveIndex"].ravel(), x=x, wx=wx)


def make_full_simulation():
    d = _mat("full_simulation_data.mat")  # tests/test_full_simulation.py:36-206
    np.savez(OUT / "full_simulation.npz", lmax=int(d["lmax"][0, 0]),
             wavelength=float(d["wavelength"][0, 0]),
             medium_refractive_index=float(d["mediumRefractiveIndex"][0, 0]),
             positions=d["particles_position"].astype(float),
             radii=d["particles_radius"].ravel().astype(float),
             refractive_index=d["particles_refractiveIndex"].ravel(),
             initial_field_coefficients=d["initial_field_coefficients"],
             mie_coefficients=d["mie_coefficients"],
             scattered_field_coefficients=d["scattered_field_coefficients"])


def _read_fits(path: Path):
    """Minimal FITS reader: headers plus raw data bytes per HDU."""
    raw = path.read_bytes()
    pos, hdus = 0, []
    while pos < len(raw):
        cards = []
        while True:
            block = raw[pos:pos + 2880]
            pos += 2880
            stop = False
            for i in range(36):
                card = block[i * 80:(i + 1) * 80].decode("ascii")
                cards.append(card)
                if card.startswith("END"):
                    stop = True
                    break
            if stop:
                break
        hdr = {}
        for c in cards:
            if "=" in c:
                k, v = c.split("=", 1)
                v = v.split("/")[0].strip().strip("'").strip()
                hdr[k.replace("HIERARCH", "").strip()] = v
        naxis = int(hdr.get("NAXIS", 0))
        size = abs(int(hdr.get("BITPIX", 8))) // 8 if naxis else 0
        for i in range(naxis):
            size *= int(hdr[f"NAXIS{i + 1}"])
        size += int(hdr.get("PCOUNT", 0))
        hdus.append((hdr, raw[pos:pos + size]))
        pos += ((size + 2879) // 2880) * 2880
    return hdus


def _table(hdr, data):
    fmt = {"D": (">f8", 1), "J": (">i4", 1), "M": (">c16", 1)}
    fields = []
    for i in range(int(hdr["TFIELDS"])):
        form = hdr[f"TFORM{i + 1}"]
        rep = int(form[:-1] or 1)
        dt, _ = fmt[form[-1]]
        fields.append((hdr[f"TTYPE{i + 1}"], dt, (rep,) if rep > 1 else ()))
    return np.frombuffer(data, dtype=np.dtype(fields), count=int(hdr["NAXIS2"]))


def make_celes():
    # tests/test_celes.py:20-150: HDU1 particles, HDU3 initial field, HDU4 input,
    # HDU6 solver, HDU7 numerics, HDU8 coefficients [re_init, im_init, re_sca, im_sca].
    (path,) = sorted(REF.glob("celes_*.fits"))
    hdus = _read_fits(path)
    part = _table(*hdus[1])
    inp = _table(*hdus[4])
    coeff = np.frombuffer(hdus[8][1], dtype=">f8").reshape(4, 500, 30).astype(float)
    np.savez(OUT / "celes500.npz",
             positions=np.stack([part["x"], part["y"], part["z"]], axis=1).astype(float),
             radii=part["r"].astype(float),
             refractive_index=(part["n"] + 1j * part["k"]).astype(complex),
             wavelength=inp["wavelength"].astype(float),
             medium_refractive_index=inp["medium_refractive_index"].astype(complex),
             lmax=int(float(hdus[7][0]["L MAX"])),
             tolerance=float(hdus[6][0]["TOLERANCE"]),
             max_iter=int(float(hdus[6][0]["MAX ITER"])),
             restart=int(float(hdus[6][0].get("RESTART", 200))),
             polar_angle=float(hdus[3][0]["POLAR ANGLE"]),
             azimuthal_angle=float(hdus[3][0]["AZIMUTHAL ANGLE"]),
             polarization=hdus[3][0]["POLARIZATION"],
             initial_field_coefficients=coeff[0] + 1j * coeff[1],
             scattered_field_coefficients=coeff[2] + 1j * coeff[3])


if __name__ == "__main__":
    for fn in (make_index, make_legendre, make_sph_bessel, make_t_entry,
               make_spherical_functions, make_transformation, make_ab5, make_coupling,
               make_full_simulation, make_celes):
        fn()
        print("wrote", fn.__name__)
