"""The literature-data loaders (src/util.jl:23-122) on a synthetic data folder
laid out like the reference's data/ (CSV columns as the reference reads
them)."""
import os

import numpy as np
import pytest


def gw():
    import geothermalwells_jl_amd as g
    return g


def _write(path, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savetxt(path, np.asarray(rows), delimiter=",")


def test_loaders(tmp_path, monkeypatch):
    d = str(tmp_path)
    _write(os.path.join(d, "Brown_et_al", "single_beier_inlet.csv"), [[10.0, 0.0], [12.0, 500.0]])
    _write(os.path.join(d, "Brown_et_al", "single_beier_outlet.csv"), [[20.0, 0.0], [18.0, 500.0]])
    _write(os.path.join(d, "Brown_et_al", "single_brown_inlet.csv"), [[11.0, 0.0], [13.0, 400.0]])
    _write(os.path.join(d, "Brown_et_al", "single_brown_outlet.csv"), [[21.0, 0.0], [19.0, 400.0]])
    (zb, Tb), (zr, Tr) = gw().data_brown_single_well_b(d)
    assert np.array_equal(zb, [0.0, 500.0, 500.0, 0.0]) and np.array_equal(Tb, [10.0, 12.0, 18.0, 20.0])
    assert np.array_equal(zr, [0.0, 400.0, 400.0, 0.0]) and np.array_equal(Tr, [11.0, 13.0, 19.0, 21.0])
    _write(os.path.join(d, "Brown_et_al", "single_600m.csv"), [[0.1, 30.0], [0.5, 31.0]])
    x, T = gw().data_brown_single_well_c(2, d)
    assert np.array_equal(x, [0.1, 0.5]) and np.array_equal(T, [30.0, 31.0])
    _write(os.path.join(d, "Hu_et_al", "Hu_5year.csv"), [[100.0, 20.0], [200.0, 25.0], [300.0, 27.0]])
    monkeypatch.setenv("GEOWELLS_DATA_DIR", d)
    z, T = gw().data_hu(5)
    assert np.array_equal(z, [100.0, 200.0, 300.0]) and np.array_equal(T, [20.0, 25.0, 27.0])
    _write(os.path.join(d, "Li_et_al", "Li_2000m.csv"), [[1.0, 70.0]])
    x, T = gw().data_li(4)
    assert x.tolist() == [1.0] and T.tolist() == [70.0]


def test_missing_folder(monkeypatch):
    monkeypatch.delenv("GEOWELLS_DATA_DIR", raising=False)
    with pytest.raises(FileNotFoundError):
        gw().data_hu(1)
