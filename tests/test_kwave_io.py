"""k-Wave HDF5 input / output (kwh_h5_*), host only: the layout the reference writes and reads
(KWaveInput.swift:12-111, HDF5.swift:26-161), checked with HDF5's own `h5dump` as an independent
reader, the known answers of the reference's IOTests.swift, and write -> read round trips."""
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from oracle import kwave_oracle as O
from kwavehip import (AbsorptionMode, KWaveGrid, KWaveMedium, KWaveSensor, KWaveSource, RecordField, makeDisc,
                      random_field, toneBurst)
from kwavehip import _lib as L
from kwavehip import kwave_io as IO

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not IO.available(), reason="libhdf5 not loadable in this environment")


def dump(path, name):
    """(HDF5 datatype, dims, values) of a dataset as h5dump prints them."""
    out = subprocess.run([H5DUMP, "-y", "-w", "0", "-m", "%.9g", "-d", name, str(path)], capture_output=True, text=True,
                         check=True).stdout
    dtype = re.search(r"DATATYPE\s+(\S+)", out).group(1)
    dims = tuple(int(x) for x in re.search(r"DATASPACE\s+SIMPLE \{ \( ([^)]*) \)", out).group(1).split(","))
    body = out[out.index("DATA {") + 6:]
    body = body[:body.index("}")]
    return dtype, dims, [float(v) for v in re.split(r"[,\s]+", body.strip()) if v]


def dump_attr(path, name, obj="/"):
    target = name if obj == "/" else f"{obj}/{name}"
    out = subprocess.run([H5DUMP, "-a", target, str(path)], capture_output=True, text=True, check=True).stdout
    return re.search(r'"([^"]*)"', out[out.index("DATA {"):]).group(1)


needs_h5dump = pytest.mark.skipif(H5DUMP is None, reason="h5dump not available")


def _small_2d(nx=3, ny=4):
    g = KWaveGrid(nx, 1e-4, ny, 1e-4)
    g.setTime(10, 2e-8)
    mask = np.zeros((nx, ny), np.float32)
    for k in range(min(nx, ny)):
        mask[k, k] = 1
    p0 = np.arange(nx * ny, dtype=np.float32).reshape(nx, ny)
    return g, KWaveMedium(1500.0, 1000.0), KWaveSource(p0=p0), KWaveSensor(mask=mask)


@needs_h5dump
def test_schema_known_answers(tmp_path):
    """IOTests.swift:23-48 and the dataset list of KWaveInput.swift:33-98."""
    path = tmp_path / "in.h5"
    g, med, src, sens = _small_2d()
    IO.writeInput2D(str(path), g, med, src, sens, pmlSize=2, pmlAlpha=2.0)
    assert dump_attr(path, "file_type") == "input"
    assert dump_attr(path, "major_version") == "1" and dump_attr(path, "minor_version") == "2"
    # field stored transposed as [1, ny, nx]: raw[i*nx + j] == field[j*ny + i]
    dtype, dims, raw = dump(path, "/p0_source_input")
    field = src.p0.reshape(-1)
    assert dtype == "H5T_IEEE_F32LE" and dims == (1, 4, 3)
    for j in range(3):
        for i in range(4):
            assert raw[i * 3 + j] == field[j * 4 + i]
    assert dump_attr(path, "domain_type", "/p0_source_input") == "real"
    assert dump_attr(path, "data_type", "/p0_source_input") == "float"
    # integer scalars as uint64 [1, 1, 1], reals as float32 [1, 1, 1]
    assert dump(path, "/Nx") == ("H5T_STD_U64LE", (1, 1, 1), [3.0])
    assert dump(path, "/Nz") == ("H5T_STD_U64LE", (1, 1, 1), [1.0])
    assert dump(path, "/Nt")[2] == [10.0]
    assert dump_attr(path, "data_type", "/Nx") == "long"
    dtype, dims, dt = dump(path, "/dt")
    assert dtype == "H5T_IEEE_F32LE" and dims == (1, 1, 1) and dt[0] == pytest.approx(2e-8, rel=1e-7)
    # sensor_mask_index: 1-based column-major indices of the diagonal mask (IOTests.swift:36)
    assert dump(path, "/sensor_mask_index") == ("H5T_STD_U64LE", (1, 1, 3), [1.0, 5.0, 9.0])
    flags = {"p0_source_flag": 1}
    for f in ("p_source_flag", "ux_source_flag", "uy_source_flag", "uz_source_flag", "sxx_source_flag",
              "syy_source_flag", "szz_source_flag", "sxy_source_flag", "sxz_source_flag", "syz_source_flag",
              "transducer_source_flag", "nonuniform_grid_flag", "nonlinear_flag", "absorbing_flag", "elastic_flag",
              "axisymmetric_flag", "sensor_mask_type", "pml_z_size"):
        flags[f] = 0
    flags.update(pml_x_size=2, pml_y_size=2)
    for f, v in flags.items():
        assert dump(path, "/" + f)[2] == [float(v)], f
    for f, v in (("c0", 1500.0), ("c_ref", 1500.0), ("rho0", 1000.0), ("rho0_sgx", 1000.0), ("rho0_sgy", 1000.0),
                 ("pml_x_alpha", 2.0), ("pml_y_alpha", 2.0)):
        assert dump(path, "/" + f)[2] == [v], f
    # 2D files carry no z variables (KWaveInput.swift:66)
    for f in ("/dz", "/pml_z_alpha", "/rho0_sgz"):
        with pytest.raises(subprocess.CalledProcessError):
            dump(path, f)


def test_reference_io_round_trip_and_missing_dataset(tmp_path):
    """IOTests.swift:5-21, 52-58 through the library's reader."""
    path = tmp_path / "rt.h5"
    g, med, src, sens = _small_2d()
    IO.writeInput(str(path), g, med, src, sens, pmlSize=2)
    raw = IO.readFloatDataset(str(path), "p0_source_input")
    assert raw.shape == (1, 4, 3)
    np.testing.assert_array_equal(raw[0].T, src.p0)
    assert IO.readRootStringAttribute(str(path), "minor_version") == "2"
    with pytest.raises(ValueError):
        IO.readFloatDataset(str(path), "nope")
    with pytest.raises(ValueError):
        IO.readRootStringAttribute(str(path), "nope")
    with pytest.raises(ValueError):
        IO.readInput(str(tmp_path / "absent.h5"))


def _hetero_3d(n=(12, 10, 8)):
    g = KWaveGrid(n[0], 1e-4, n[1], 2e-4, n[2], 1.5e-4)
    g.setTime(17, 1.5e-8)
    c = random_field(n, 5, 2.0, 1400.0, 1600.0)
    rho = random_field(n, 6, 2.0, 900.0, 1100.0)
    bona = random_field(n, 7, 2.0, 4.0, 8.0)
    mask = np.zeros(n, np.float32)
    mask[2, 3:7, 1] = 1
    mask[9, 0, 0:8:3] = 1
    mask[5, 9, 7] = 1
    p0 = random_field(n, 8, 1.5, 0.0, 1.0)
    return g, KWaveMedium(c, rho, bOnA=bona), KWaveSource(p0=p0), KWaveSensor(mask=mask)


def test_round_trip_3d_maps_and_nonlinearity(tmp_path):
    path = tmp_path / "in3d.h5"
    g, med, src, sens = _hetero_3d()
    IO.writeInput(str(path), g, med, src, sens, pmlSize=3, pmlAlpha=1.5)
    g2, med2, src2, sens2, opts2 = IO.readInput(str(path))
    assert g2.dim == 3 and list(g2.size) == list(g.size) and g2.nt == g.nt
    for a in range(3):
        assert g2.spacing[a] == float(np.float32(g.spacing[a]))
    assert g2.dt == float(np.float32(g.dt))
    np.testing.assert_array_equal(med2.soundSpeed, med.soundSpeed)
    np.testing.assert_array_equal(med2.density, med.density)
    np.testing.assert_array_equal(med2.bOnA, med.bOnA)
    np.testing.assert_array_equal(src2.p0, src.p0)
    np.testing.assert_array_equal(sens2.mask, sens.mask)
    assert opts2.pmlSize == 3 and opts2.pmlAlpha == 1.5 and opts2.smoothP0 is False
    assert IO.readFloatDataset(str(path), "nonlinear_flag").item() == 1.0
    # staggered densities: the neighbour mean along each axis, the last cell replicated
    rho = med.density.astype(np.float64)
    for a, name in enumerate(("rho0_sgx", "rho0_sgy", "rho0_sgz")):
        disk = IO.readFloatDataset(str(path), name)
        want = O.stagger_density(rho, a).astype(np.float32)
        np.testing.assert_allclose(disk.transpose(2, 1, 0), want, rtol=1e-7)
    # c_ref = max(c0)
    assert IO.readFloatDataset(str(path), "c_ref").item() == float(med.soundSpeed.max())


@needs_h5dump
def test_3d_field_layout_and_sensor_order(tmp_path):
    path = tmp_path / "layout3d.h5"
    g, med, src, sens = _hetero_3d()
    IO.writeInput(str(path), g, med, src, sens, pmlSize=3)
    nx, ny, nz = g.size
    _, dims, raw = dump(path, "/c0")
    assert dims == (nz, ny, nx)
    raw = np.asarray(raw, np.float32).reshape(nz, ny, nx)   # %.9g round-trips float32 exactly
    c = med.soundSpeed
    for x, y, z in ((0, 0, 0), (11, 0, 0), (0, 9, 0), (0, 0, 7), (3, 5, 6)):
        assert raw[z, y, x] == c[x, y, z]
    _, _, idx = dump(path, "/sensor_mask_index")
    pts = np.argwhere(sens.mask > 0)                       # C-order (x, y, z)
    want = sorted(int(x + nx * (y + ny * z)) + 1 for x, y, z in pts)
    assert idx == [float(v) for v in want]
    assert dump(path, "/Nz")[2] == [float(nz)] and dump(path, "/pml_z_size")[2] == [3.0]
    assert len(dump(path, "/dz")[2]) == 1


def test_unsupported_inputs_are_rejected(tmp_path):
    g, med, src, sens = _small_2d(8, 8)
    absorbing = KWaveMedium(1500.0, 1000.0, alphaCoeff=0.5, alphaPower=1.5, alphaMode=AbsorptionMode.powerLaw)
    with pytest.raises(NotImplementedError):
        IO.writeInput(str(tmp_path / "a.h5"), g, absorbing, src, sens)
    cart = KWaveSensor(mask=np.array([[0.0, 1e-4], [0.0, 0.0]], np.float32))
    with pytest.raises(NotImplementedError):
        IO.writeInput(str(tmp_path / "b.h5"), g, med, src, cart)
    mask = np.zeros((8, 8), np.float32)
    mask[4, 4] = 1
    tv = KWaveSource(p0=src.p0, pMask=mask, p=toneBurst(1 / g.dt, 1e6, 2).astype(np.float32))
    with pytest.raises(NotImplementedError):
        IO.writeInput(str(tmp_path / "c.h5"), g, med, tv, sens)
    g1 = KWaveGrid(16, 1e-4)
    g1.setTime(5, 2e-8)
    with pytest.raises(NotImplementedError):
        IO.writeInput(str(tmp_path / "d.h5"), g1, med, KWaveSource(p0=np.ones(16, np.float32)),
                      KWaveSensor(mask=np.ones(16, np.float32)))


def test_library_does_not_link_hdf5():
    out = subprocess.run(["readelf", "-d", L.LIB_PATH], capture_output=True, text=True, check=True).stdout
    needed = re.findall(r"\(NEEDED\).*\[(.*)\]", out)
    assert needed and not any("hdf5" in n for n in needed), needed
