"""Cluster geometry: the fractal-aggregate generator and the cluster-file reader
(CPU).  The two PyFracVAL files are data shipped with the reference's examples
(examples/spaceweather/pyfracval/), used as input fixtures."""

from pathlib import Path

import numpy as np
import pytest
from scipy.spatial import cKDTree

from yasf_amd.aggregates import fractal_aggregate, radius_of_gyration, read_cluster

GOLDEN = Path(__file__).parent / "golden"


def _contacts(pos, a):
    tree = cKDTree(pos)
    pairs = tree.query_pairs(2 * a * (1 + 1e-9))
    d = np.array([np.linalg.norm(pos[i] - pos[j]) for i, j in pairs])
    return pairs, d


@pytest.mark.parametrize("N,Df,kf", [(64, 1.8, 1.3), (500, 1.8, 1.3), (300, 2.2, 1.2)])
def test_fractal_law_contacts_and_connectivity(N, Df, kf):
    a = 40.0
    pos, radii = fractal_aggregate(11, N, Df, kf, a)
    assert pos.shape == (N, 3) and np.all(radii == a)
    # the fractal law N = kf (Rg/a)^Df of the construction, within 10%
    law = kf * (radius_of_gyration(pos) / a) ** Df
    assert abs(N / law - 1) < 0.1
    # no overlaps; every contact at exactly 2a
    d, _ = cKDTree(pos).query(pos, k=2)
    assert d[:, 1].min() >= 2 * a * (1 - 1e-9)
    pairs, dc = _contacts(pos, a)
    assert np.allclose(dc, 2 * a, rtol=1e-9)
    # one connected aggregate (union-find over contacts)
    parent = list(range(N))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    for i, j in pairs:
        parent[find(i)] = find(j)
    assert len({find(i) for i in range(N)}) == 1
    # centred on the centre of mass; seeded
    assert np.allclose(pos.mean(axis=0), 0, atol=1e-9 * a * N)
    pos2, _ = fractal_aggregate(11, N, Df, kf, a)
    assert np.array_equal(pos, pos2)


def test_read_pyfracval_files():
    for name, N, Df in (("fracval_N8_Df1p80.dat", 8, 1.8), ("fracval_N48_Df2p00.dat", 48, 2.0)):
        spheres, meta = read_cluster(GOLDEN / name)
        assert spheres.shape == (N, 4)
        assert meta["simulation_parameters"]["N"] == N
        assert meta["simulation_parameters"]["Df"] == pytest.approx(Df)
        assert meta["aggregate_properties"]["N_particles_actual"] == N
        assert np.all(spheres[:, 3] > 0)


def test_read_csv_and_delimited(tmp_path):
    data = np.array([[0.0, 0.0, 0.0, 50.0, 1.5], [200.0, 0.0, 0.0, 60.0, 1.6]])
    csv = tmp_path / "c.csv"
    csv.write_text("x,y,z,r,n\n" + "\n".join(",".join(map(str, r)) for r in data))
    spheres, meta = read_cluster(csv)
    assert np.array_equal(spheres, data) and meta is None
    txt = tmp_path / "c.txt"
    txt.write_text("\n".join(" ".join(map(str, r)) for r in data))
    assert np.array_equal(read_cluster(txt, delimiter="whitespace")[0], data)
    with pytest.raises(NotImplementedError):
        read_cluster(txt)
    bad = tmp_path / "b.csv"
    bad.write_text("x,y,z\n1,2,3\n")
    with pytest.raises(ValueError):
        read_cluster(bad)
